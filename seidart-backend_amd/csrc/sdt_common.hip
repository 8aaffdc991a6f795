// sdt_common.hip -- per-step tail kernel (point sources, norm fix-up,
// receivers, step counter) and setup helpers shared by all solvers.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <unordered_map>
#include <vector>

#include "sdt_device.h"
#include "sdt_internal.h"

namespace sdt {

template <typename T>
__device__ __forceinline__ T *fptr(void *const *field, int f)
{
    return reinterpret_cast<T *>(field[f]);
}

struct FieldPtrs {
    void *f[SDT_NUM_FIELDS];
};

// Index of cell c in the source window's value arrays, -1 outside the window.
__device__ __forceinline__ int win_index(const SrcWindow &w, int64_t c)
{
    if (!w.v || c < w.base) return -1;
    const int64_t o = c - w.base, jj = o / w.pitch, ii = o - jj * w.pitch;
    return ii < w.ni && jj < w.nj ? (int)(ii + (int64_t)(w.ni + 1) * jj) : -1;
}

__device__ __forceinline__ const double *win_values(const SrcWindow &w, int f)
{
    const int m = (w.ni + 1) * (w.nj + 1);
    return f == SDT_F_EX ? w.v : f == SDT_F_EZ ? w.v + m : f == SDT_F_HY ? w.v + 2 * m : nullptr;
}

// The FP64 source window of a 2D EM shot (whole workgroup, after the step's
// FP32 H and E kernels): the Hy update of cpmlfdtd.f08:2576-2593 and the Ex,
// Ez update of 2605-2620 for the window's cells, without CPML terms (the
// window lies in the CPML-free box).  The H update reads the window's FP64 E
// and, past its high edges, the ring (the FP32 E the previous step left
// there); the E update reads the window's new FP64 Hy and, below its low
// edges, the FP32 Hy the H kernel just wrote.  The point-source terms on a
// window cell are added in FP64 (2678-2681); the rounded values overwrite
// the FP32 fields, and the ring is refreshed from the FP32 E of this step.
template <typename T>
__device__ void src_window_step(const FieldPtrs &fp, const TailArgs &a, int it)
{
    const SrcWindow &w = a.win;
    const int ni = w.ni, nj = w.nj, m = ni + 1, t = threadIdx.x;
    double *vx = w.v, *vz = w.v + m * (nj + 1), *vh = w.v + 2 * m * (nj + 1);
    T *Ex = fptr<T>(fp.f, SDT_F_EX), *Ez = fptr<T>(fp.f, SDT_F_EZ), *Hy = fptr<T>(fp.f, SDT_F_HY);
    const bool on = t < ni * nj;
    const int ii = on ? t % ni : 0, jj = on ? t / ni : 0, q = ii + m * jj;
    const int64_t c = w.base + ii + (int64_t)w.pitch * jj;
    double h = 0.0;
    if (on) {
        const double dEx_dz = (vx[q + m] - vx[q]) * w.rdz;
        const double dEz_dx = (vz[q + 1] - vz[q]) * w.rdx;
        h = vh[q] + w.dbHy * (dEz_dx - dEx_dz);
    }
    __syncthreads();
    if (on) vh[q] = h;
    __syncthreads();
    if (on) {
        const double hd = jj > 0 ? vh[q - m] : (double)Hy[c - w.pitch];
        const double hl = ii > 0 ? vh[q - 1] : (double)Hy[c - 1];
        const double exo = vx[q], ezo = vz[q];
        const int nc = ni * nj;
        const double aEx = w.co[t], bEx = w.co[nc + t], bEz = w.co[2 * nc + t], dEx = w.co[3 * nc + t],
                     eEx = w.co[4 * nc + t], dEz = w.co[5 * nc + t], eEz = w.co[6 * nc + t];
        const double dHy_dz = (h - hd) * w.rdz;
        const double dHy_dx = (h - hl) * w.rdx;
        double ez = ezo + (-bEx * dHy_dz + bEz * dHy_dx - dEz * exo - eEz * ezo) * w.dt;
        double ex = exo + (-aEx * dHy_dz + bEx * dHy_dx - dEx * exo - eEx * ezo) * w.dt;
        for (int k = 0; k < a.nterms; ++k) {
            const SrcTerm &s = a.terms[k];
            if (s.cell != c || s.on_face) continue;
            const double v = (a.series[(int64_t)s.component * a.time_steps + (it - 1)] * s.mul) / s.div;
            if (s.field == SDT_F_EX) ex += v;
            else if (s.field == SDT_F_EZ) ez += v;
        }
        vx[q] = ex;
        vz[q] = ez;
        Ex[c] = (T)ex;
        Ez[c] = (T)ez;
        Hy[c] = (T)h;
    }
    if (t < ni + nj) {   // ring column (t < nj), ring row
        const int ri = t < nj ? ni : t - nj, rj = t < nj ? t : nj;
        const int64_t rc = w.base + ri + (int64_t)w.pitch * rj;
        vx[ri + m * rj] = (double)Ex[rc];
        vz[ri + m * rj] = (double)Ez[rc];
    }
    __syncthreads();
}

// One workgroup.  The point sources are injected in parallel, one thread per
// term: every term of a solver adds to a different (field, cell) pair, so the
// reference's sequential order (e.g. cpmlfdtd.f08:519-523, 2208-2219,
// 2678-2681, 3909-3914) gives the same values.  After a barrier the source
// cells' norms are folded in and the receivers recorded; then thread 0
// advances the device step counter.
template <typename T>
__device__ __forceinline__ void tail_body(const FieldPtrs &fp, const TailArgs &a)
{
    const int it = *a.d_it;
    const int t = threadIdx.x;
    if (a.win.v) src_window_step<T>(fp, a, it);   // uniform per workgroup
    if (t < a.nterms) {
        const SrcTerm &s = a.terms[t];
        if (s.cell >= 0 && !s.on_face && win_index(a.win, s.cell) < 0) {
            const double v = a.series[(int64_t)s.component * a.time_steps + (it - 1)];
            T *f = reinterpret_cast<T *>(fp.f[s.field]);
            const int64_t e = s.cell << a.cplx;   // complex: the real part
            f[e] = (T)((double)f[e] + (T)((v * s.mul) / s.div));
        }
    }
    __syncthreads();
    if (t < a.nvfix) {
        double ss = 0.0;
        for (int c = 0; c < a.nvcomp; ++c) {
            const T *f = reinterpret_cast<const T *>(fp.f[a.vcomp[c]]);
            if (a.cplx) {
                const double re = (double)f[2 * a.vfix[t]], im = (double)f[2 * a.vfix[t] + 1];
                ss += re * re + im * im;
            } else {
                const double v = (double)f[a.vfix[t]];
                ss += v * v;
            }
        }
        atomic_max_nonneg(&a.vnorm[it - 1], sqrt(ss));
    } else if (a.stress_norm && t >= 64 && t < 64 + a.nsfix) {
        const int64_t c = a.sfix[t - 64];
        const double sxx = (double)reinterpret_cast<T *>(fp.f[SDT_F_SXX])[c];
        const double syy = (double)reinterpret_cast<T *>(fp.f[SDT_F_SYY])[c];
        const double szz = (double)reinterpret_cast<T *>(fp.f[SDT_F_SZZ])[c];
        const double sxy = (double)reinterpret_cast<T *>(fp.f[SDT_F_SXY])[c];
        const double sxz = (double)reinterpret_cast<T *>(fp.f[SDT_F_SXZ])[c];
        const double syz = (double)reinterpret_cast<T *>(fp.f[SDT_F_SYZ])[c];
        const double n = sqrt(sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + syz * syz + sxz * sxz));
        atomic_max_nonneg(&a.snorm[it - 1], n);
    }
    for (int q = t; q < a.nrec; q += blockDim.x) {
        const int64_t c = a.rec_cell[q];
        const int nv = 1 << a.cplx;
        double *o = a.rec_out + ((int64_t)(it - 1) * a.nrec + q) * 3 * nv;
        const int wq = win_index(a.win, c);   // inside the source window: its FP64 values
        for (int m = 0; m < 3; ++m)
            for (int p = 0; p < nv; ++p)
                o[m * nv + p] = a.rec_field[m] < 0 ? 0.0
                                : wq >= 0          ? win_values(a.win, a.rec_field[m])[wq]
                                                   : (double)reinterpret_cast<T *>(fp.f[a.rec_field[m]])[(c << a.cplx) + p];
    }
    __syncthreads();
    if (t == 0) *a.d_it = it + 1;
}

template <typename T>
__global__ void __launch_bounds__(256) tail_kernel(FieldPtrs fp, TailArgs a)
{
    tail_body<T>(fp, a);
}

// one workgroup per shot of a batch; finished shots do nothing
template <typename T>
__global__ void __launch_bounds__(256) tail_batch_kernel(const FieldPtrs *fps, const TailArgs *tails)
{
    const TailArgs &a = tails[blockIdx.x];
    if (*a.d_it > a.time_steps) return;
    tail_body<T>(fps[blockIdx.x], a);
}

void launch_tail_batch(Batch &b)
{
    if (b.precision == SDT_FP64)
        hipLaunchKernelGGL(tail_batch_kernel<double>, dim3(b.nshots), dim3(256), 0, b.stream,
                           (const FieldPtrs *)b.fps, (const TailArgs *)b.tails);
    else
        hipLaunchKernelGGL(tail_batch_kernel<float>, dim3(b.nshots), dim3(256), 0, b.stream,
                           (const FieldPtrs *)b.fps, (const TailArgs *)b.tails);
}

int batch_setup(Batch &b, Device *const *shots, int n)
{
    b.nshots = n;
    b.physics = shots[0]->physics;
    b.precision = shots[0]->precision;
    b.stream = shots[0]->stream;
    // tile lists for every candidate tile height
    for (int w = 0; w < kNumBatchShapes; ++w) {
        std::vector<host::BatchTile> t;
        for (int s = 0; s < n; ++s) {
            const int tx = tiles_x_2d(*shots[s]), tz = tiles_z_2d(*shots[s], kBatchShapes[w].rows);
            for (int z = 0; z < tz; ++z)
                for (int x = 0; x < tx; ++x) t.push_back(host::BatchTile{s, x, z, 0});
        }
        if (kBatchShapes[w].xcd_order) t = host::batch_order(t);
        static_assert(sizeof(host::BatchTile) == sizeof(int4), "batch tiles are read as int4");
        b.ntiles[w] = (int)t.size();
        if (hipMalloc((void **)&b.tiles[w], t.size() * sizeof(int4)) != hipSuccess) return SDT_ERR_HIP;
        if (hipMemcpy(b.tiles[w], t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess)
            return SDT_ERR_HIP;
    }
    std::vector<FieldPtrs> fps((size_t)n);
    std::vector<TailArgs> tails((size_t)n);
    for (int s = 0; s < n; ++s) {
        for (int f = 0; f < SDT_NUM_FIELDS; ++f) fps[(size_t)s].f[f] = shots[s]->field[f];
        tails[(size_t)s] = shots[s]->tail;
    }
    if (hipMalloc(&b.fps, fps.size() * sizeof(FieldPtrs)) != hipSuccess) return SDT_ERR_HIP;
    if (hipMemcpy(b.fps, fps.data(), fps.size() * sizeof(FieldPtrs), hipMemcpyHostToDevice) != hipSuccess)
        return SDT_ERR_HIP;
    if (hipMalloc((void **)&b.tails, tails.size() * sizeof(TailArgs)) != hipSuccess) return SDT_ERR_HIP;
    if (hipMemcpy(b.tails, tails.data(), tails.size() * sizeof(TailArgs), hipMemcpyHostToDevice) != hipSuccess)
        return SDT_ERR_HIP;
    if (b.physics == SDT_ELASTIC2) return el2_batch_args(b, shots, n);
    if (b.physics == SDT_EM2) return em2_batch_args(b, shots, n);
    return SDT_ERR_ARG;
}

void batch_free(Batch &b)
{
    for (int w = 0; w < kNumBatchShapes; ++w)
        if (b.tiles[w]) (void)hipFree(b.tiles[w]);
    if (b.fps) (void)hipFree(b.fps);
    if (b.tails) (void)hipFree(b.tails);
    if (b.args) (void)hipFree(b.args);
    if (b.adj_args) (void)hipFree(b.adj_args);
    if (b.adj_tails) (void)hipFree(b.adj_tails);
    b = Batch();
}

void batch_step(Batch &b, int which)
{
    if (b.physics == SDT_ELASTIC2) el2_batch_launch(b, which);
    else em2_batch_launch(b, which);
    launch_tail_batch(b);
}

// ---- adjoint helpers ---------------------------------------------------------
// Trace-gradient row t added to the adjoint of the recorded fields at the
// receiver cells (atomics: receivers may share a cell).
template <typename T>
__device__ __forceinline__ void adj_tail_body(const AdjTailArgs &a, int t)
{
    for (int q = threadIdx.x; q < a.nrec; q += blockDim.x)
        for (int m = 0; m < 3; ++m)
            if (a.lam[m])
                atomicAdd(reinterpret_cast<T *>(a.lam[m]) + a.rec_cell[q],
                          (T)a.tg[((int64_t)t * a.nrec + q) * 3 + m]);
}

template <typename T>
__global__ void __launch_bounds__(256) adj_tail_kernel(AdjTailArgs a, int t)
{
    adj_tail_body<T>(a, t);
}

// one workgroup per shot
template <typename T>
__global__ void __launch_bounds__(256) adj_tail_batch_kernel(const AdjTailArgs *a, int t)
{
    adj_tail_body<T>(a[blockIdx.x], t);
}

void launch_adj_tail_batch(Batch &b, int t)
{
    if (b.precision == SDT_FP64)
        hipLaunchKernelGGL(adj_tail_batch_kernel<double>, dim3(b.nshots), dim3(256), 0, b.stream, b.adj_tails, t);
    else
        hipLaunchKernelGGL(adj_tail_batch_kernel<float>, dim3(b.nshots), dim3(256), 0, b.stream, b.adj_tails, t);
}

int batch_refresh_args(Batch &b, Device *const *shots, int n)
{
    if (hipStreamSynchronize(b.stream) != hipSuccess) return SDT_ERR_HIP;   // no launch reads them now
    if (b.physics == SDT_ELASTIC2) return el2_batch_args(b, shots, n);
    return em2_batch_args(b, shots, n);
}

int batch_adj_args(Batch &b, Device *const *shots, int n)
{
    if (hipStreamSynchronize(b.stream) != hipSuccess) return SDT_ERR_HIP;
    if (b.physics == SDT_ELASTIC2) return el2_batch_adj_args(b, shots, n);
    return em2_batch_adj_args(b, shots, n);
}

void batch_adj_step(Batch &b, int which, int t, int first)
{
    launch_adj_tail_batch(b, t);
    if (b.physics == SDT_ELASTIC2) el2_batch_adj_step(b, which, t, first);
    else em2_batch_adj_step(b, which, t, first);
}

void batch_adj_correlate(Batch &b, int which, int n)
{
    if (b.physics == SDT_ELASTIC2) el2_batch_adj_correlate(b, which, n);
    else em2_batch_adj_correlate(b, which, n);
}

void launch_adj_tail(Device &d, int t, void *const lam_field[3])
{
    if (d.nrec == 0) return;
    AdjTailArgs a;
    for (int m = 0; m < 3; ++m) a.lam[m] = lam_field[m];
    a.rec_cell = d.rec_cell;
    a.tg = d.adj.trace_grad;
    a.nrec = d.nrec;
    if (d.precision == SDT_FP64)
        hipLaunchKernelGGL(adj_tail_kernel<double>, dim3(1), dim3(256), 0, d.stream, a, t);
    else
        hipLaunchKernelGGL(adj_tail_kernel<float>, dim3(1), dim3(256), 0, d.stream, a, t);
}

__global__ void set_counter_kernel(int *p, int v)
{
    if (threadIdx.x == 0) *p = v;
}

void set_step_counter(hipStream_t s, int *d_it, int v)
{
    hipLaunchKernelGGL(set_counter_kernel, dim3(1), dim3(64), 0, s, d_it, v);
}


void launch_tail(Device &d)
{
    FieldPtrs fp;
    for (int f = 0; f < SDT_NUM_FIELDS; ++f) fp.f[f] = d.field[f];
    if (d.precision == SDT_FP64)
        hipLaunchKernelGGL(tail_kernel<double>, dim3(1), dim3(256), 0, d.stream, fp, d.tail);
    else
        hipLaunchKernelGGL(tail_kernel<float>, dim3(1), dim3(256), 0, d.stream, fp, d.tail);
}

int alloc_grid(Device &d, void **dev, size_t bytes)
{
    const size_t shift = d.stagger ? (size_t)(d.nalloc % 61) * 4352u : 0;
    d.nalloc++;
    if (d.nbases >= (int)(sizeof(d.bases) / sizeof(d.bases[0]))) return SDT_ERR_HIP;
    void *base = nullptr;
    if (hipMalloc(&base, bytes + shift) != hipSuccess) return SDT_ERR_HIP;
    d.bases[d.nbases++] = base;
    d.grid_bytes += bytes + shift;
    *dev = (char *)base + shift;
    return SDT_OK;
}

void free_grids(Device &d)
{
    for (int q = 0; q < d.nbases; ++q)
        if (d.bases[q]) (void)hipFree(d.bases[q]);
    d.nbases = 0;
}

int upload_converted(Device &d, const double *host, void **dev, int64_t n)
{
    const size_t bytes = (size_t)n * d.esize;
    if (alloc_grid(d, dev, bytes) != SDT_OK) return SDT_ERR_HIP;
    if (d.precision == SDT_FP64) {
        if (hipMemcpy(*dev, host, bytes, hipMemcpyHostToDevice) != hipSuccess) return SDT_ERR_HIP;
    } else {
        std::vector<float> tmp((size_t)n);
        for (int64_t i = 0; i < n; ++i) tmp[(size_t)i] = (float)host[i];
        if (hipMemcpy(*dev, tmp.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return SDT_ERR_HIP;
    }
    return SDT_OK;
}

int alloc_zeroed(Device &d, void **dev, int64_t n)
{
    const size_t bytes = (size_t)n * d.esize;
    if (alloc_grid(d, dev, bytes) != SDT_OK) return SDT_ERR_HIP;
    if (hipMemset(*dev, 0, bytes) != hipSuccess) return SDT_ERR_HIP;
    return SDT_OK;
}

int upload_cpml_dict(Device &d, const double *const *in)
{
    const int cids[6] = {SDT_ACOEF, SDT_BCOEF, SDT_KAPPA, SDT_ACOEF_HALF, SDT_BCOEF_HALF, SDT_KAPPA_HALF};
    const host::RowDictionary cd = host::dedupe_rows(d.ncell, 8, [&](int64_t c, double *r) {
        for (int q = 0; q < 6; ++q) {
            const double *src = in[cids[q]];
            const bool inverse = (q == 2 || q == 5);
            const double v = src ? src[c] : (inverse ? 1.0 : 0.0);
            r[q] = inverse ? 1.0 / v : v;
        }
    });
    if (!cd.ok) return upload_cpml_full(d, in);
    const size_t bytes = (size_t)d.ncell_dev;
    if (hipMalloc((void **)&d.cpml_cid, bytes) != hipSuccess) return SDT_ERR_HIP;
    if (hipMemset(d.cpml_cid, 0, bytes) != hipSuccess) return SDT_ERR_HIP;
    if (hipMemcpy2D(d.cpml_cid + d.xoff, (size_t)d.pitch, cd.ids.data(), (size_t)d.nx, (size_t)d.nx,
                    (size_t)d.ny * d.nz, hipMemcpyHostToDevice) != hipSuccess)
        return SDT_ERR_HIP;
    if (upload_converted(d, cd.rows.data(), &d.cpml_ctab, (int64_t)cd.rows.size()) != SDT_OK) return SDT_ERR_HIP;
    d.cpml_ncoef = cd.count();
    return SDT_OK;
}

int upload_cpml_full(Device &d, const double *const *in)
{
    const int cids[6] = {SDT_ACOEF, SDT_BCOEF, SDT_KAPPA, SDT_ACOEF_HALF, SDT_BCOEF_HALF, SDT_KAPPA_HALF};
    std::vector<double> tmp((size_t)d.ncell);
    for (int q = 0; q < 6; ++q) {
        const double *src = in[cids[q]];
        const bool inverse = (q == 2 || q == 5);
        for (int64_t c = 0; c < d.ncell; ++c) {
            const double v = src ? src[c] : (inverse ? 1.0 : 0.0);
            tmp[(size_t)c] = inverse ? 1.0 / v : v;
        }
        int rc = upload_grid(d, tmp.data(), &d.cpml[q]);
        if (rc) return rc;
    }
    return SDT_OK;
}

void set_layout(Device &d, int first)
{
    const bool empty = d.box[1] < d.box[0];
    if (first <= 0) first = empty ? 1 : d.box[0];
    const host::RowLayout r = host::row_layout(d.nx, first, (int)d.esize);
    d.xoff = r.xoff;
    d.pitch = r.pitch;
    d.ncell_dev = (int64_t)d.pitch * d.ny * d.nz;
}

// rows of the grid = ny * nz, each nx elements on the host, pitch on the
// device; an element is ncomp values (2 = complex, interleaved re/im)
int upload_grid(Device &d, const double *host, void **dev, int ncomp)
{
    const size_t bytes = (size_t)d.ncell_dev * d.esize * ncomp;
    if (alloc_grid(d, dev, bytes) != SDT_OK) return SDT_ERR_HIP;
    return store_grid(d, *dev, host, ncomp);   // writes the padding too
}

int alloc_grid_zeroed(Device &d, void **dev, int ncomp)
{
    const size_t bytes = (size_t)d.ncell_dev * d.esize * ncomp;
    if (alloc_grid(d, dev, bytes) != SDT_OK) return SDT_ERR_HIP;
    if (hipMemset(*dev, 0, bytes) != hipSuccess) return SDT_ERR_HIP;
    return SDT_OK;
}

// Host <-> device copies of whole grid arrays go through one contiguous
// host image of the padded layout (zero padding), converted to the device
// precision on the host: one copy per array instead of one per row.
template <typename T>
static int store_grid_t(Device &d, void *dev, const double *host, int ncomp)
{
    std::vector<T> img;
    host::pack_rows<T>(host, (size_t)d.ny * d.nz, (size_t)d.nx, (size_t)d.pitch, (size_t)d.xoff, ncomp, img);
    return hipMemcpy(dev, img.data(), img.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess ? SDT_OK
                                                                                                 : SDT_ERR_HIP;
}

template <typename T>
static int download_grid_t(Device &d, const void *dev, double *host, int ncomp)
{
    const size_t rows = (size_t)d.ny * d.nz;
    std::vector<T> img(rows * d.pitch * ncomp);
    if (hipMemcpy(img.data(), dev, img.size() * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) return SDT_ERR_HIP;
    host::unpack_rows<T>(img, rows, (size_t)d.nx, (size_t)d.pitch, (size_t)d.xoff, ncomp, host);
    return SDT_OK;
}

int store_grid(Device &d, void *dev, const double *host, int ncomp)
{
    return d.precision == SDT_FP64 ? store_grid_t<double>(d, dev, host, ncomp) : store_grid_t<float>(d, dev, host, ncomp);
}

int download_grid(Device &d, const void *dev, double *host, int ncomp)
{
    return d.precision == SDT_FP64 ? download_grid_t<double>(d, dev, host, ncomp)
                                   : download_grid_t<float>(d, dev, host, ncomp);
}

// an FP64 array in the device grid layout of d -> host (nx, ny, nz) order
int download_grid_f64(Device &d, const double *dev, double *host)
{
    return download_grid_t<double>(d, dev, host, 1);
}

// ---- snapshots ---------------------------------------------------------------
// Planes k0 .. k0+nk-1 (0-based) of a grid array in the device layout (rows
// of `pitch` elements, x cells from xoff) -> Fortran order (x fastest, no
// padding) in dst, converted to the output type.  With zero_faces the cells
// on the x and z faces (and the y faces when ny > 1) read 0: the Dirichlet
// faces of H that the EM kernels apply lazily (write_image2/3 of
// seidartio.f08:366-426 write the field after the Dirichlet step).
template <typename TS, typename TD>
__global__ void __launch_bounds__(256) pack_kernel(const TS *__restrict__ src, TD *__restrict__ dst, int nx, int ny,
                                                   int nz, int k0, int nk, int pitch, int64_t xoff, int zero_faces)
{
    const int64_t n = (int64_t)nx * ny * nk;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int i = (int)(e % nx);
    const int64_t r = e / nx;
    const int j = (int)(r % ny), k = k0 + (int)(r / ny);
    TS v = src[xoff + i + (int64_t)pitch * (j + (int64_t)ny * k)];
    if (zero_faces && (i == 0 || i == nx - 1 || k == 0 || k == nz - 1 || (ny > 1 && (j == 0 || j == ny - 1))))
        v = TS(0);
    dst[e] = (TD)v;
}

void launch_pack(const Device &d, const void *src, void *dst, int single, int k0, int nk, bool zero_faces,
                 hipStream_t s)
{
    const int64_t n = (int64_t)d.nx * d.ny * nk;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    const int zf = zero_faces ? 1 : 0;
    if (d.precision == SDT_FP64) {
        if (single)
            hipLaunchKernelGGL((pack_kernel<double, float>), grid, block, 0, s, (const double *)src, (float *)dst,
                               d.nx, d.ny, d.nz, k0, nk, d.pitch, d.xoff, zf);
        else
            hipLaunchKernelGGL((pack_kernel<double, double>), grid, block, 0, s, (const double *)src,
                               (double *)dst, d.nx, d.ny, d.nz, k0, nk, d.pitch, d.xoff, zf);
    } else {
        if (single)
            hipLaunchKernelGGL((pack_kernel<float, float>), grid, block, 0, s, (const float *)src, (float *)dst,
                               d.nx, d.ny, d.nz, k0, nk, d.pitch, d.xoff, zf);
        else
            hipLaunchKernelGGL((pack_kernel<float, double>), grid, block, 0, s, (const float *)src, (double *)dst,
                               d.nx, d.ny, d.nz, k0, nk, d.pitch, d.xoff, zf);
    }
}

int build_material_table(Device &d, int row_len, const std::function<void(int64_t, double *)> &fill)
{
    const host::RowDictionary md = host::dedupe_rows(d.ncell, row_len, fill);
    if (!md.ok) return SDT_OK;   // too many materials: per-cell arrays
    const std::vector<uint8_t> &ids = md.ids;
    const std::vector<double> &rows = md.rows;
    // ids in the padded layout (1 byte per element)
    const size_t bytes = (size_t)d.ncell_dev;
    if (hipMalloc((void **)&d.mid, bytes) != hipSuccess) return SDT_ERR_HIP;
    if (hipMemset(d.mid, 0, bytes) != hipSuccess) return SDT_ERR_HIP;
    if (hipMemcpy2D(d.mid + d.xoff, (size_t)d.pitch, ids.data(), (size_t)d.nx, (size_t)d.nx,
                    (size_t)d.ny * d.nz, hipMemcpyHostToDevice) != hipSuccess)
        return SDT_ERR_HIP;
    if (upload_converted(d, rows.data(), &d.tab, (int64_t)rows.size()) != SDT_OK) return SDT_ERR_HIP;
    d.nmaterials = (int)(rows.size() / row_len);
    d.mat_table = true;
    return SDT_OK;
}

int compute_tile_flags(Device &d, const double *const *host_in)
{
    const host::BoxScan bs = host::scan_box(d.nx, d.ny, d.nz, d.tile_x, d.tile_y, d.tile_z, host_in[SDT_ACOEF],
                                            host_in[SDT_BCOEF], host_in[SDT_KAPPA], host_in[SDT_ACOEF_HALF],
                                            host_in[SDT_BCOEF_HALF], host_in[SDT_KAPPA_HALF]);
    d.tiles_x = bs.tiles_x;
    d.tiles_y = bs.tiles_y;
    d.tiles_z = bs.tiles_z;
    for (int q = 0; q < 6; ++q) d.box[q] = bs.box[q];
    d.nflagged += bs.nflagged;
    if (hipMalloc((void **)&d.tile_flags, bs.flags.size()) != hipSuccess) return SDT_ERR_HIP;
    if (hipMemcpy(d.tile_flags, bs.flags.data(), bs.flags.size(), hipMemcpyHostToDevice) != hipSuccess)
        return SDT_ERR_HIP;
    return SDT_OK;
}

// Isotropic form of every cell's stiffness: c11 = c22 = c33, c12 = c13 = c23,
// c44 = c55 = c66 and the other twelve entries 0 (exact comparisons; absent
// arrays are 0).  The stress updates then read c11, c12, c44 only, with the
// same products and sums as from the full tensor (the zero entries add exact
// zeros).
bool stiffness_isotropic(const double *const *in, int64_t n)
{
    auto at = [&](int id, int64_t c) { return in[id] ? in[id][c] : 0.0; };
    static const int zero_ids[] = {SDT_C14, SDT_C15, SDT_C16, SDT_C24, SDT_C25, SDT_C26,
                                   SDT_C34, SDT_C35, SDT_C36, SDT_C45, SDT_C46, SDT_C56};
    for (int64_t c = 0; c < n; ++c) {
        const double a = at(SDT_C11, c), l = at(SDT_C12, c), m = at(SDT_C44, c);
        if (at(SDT_C22, c) != a || at(SDT_C33, c) != a || at(SDT_C13, c) != l || at(SDT_C23, c) != l ||
            at(SDT_C55, c) != m || at(SDT_C66, c) != m)
            return false;
        for (int id : zero_ids)
            if (at(id, c) != 0.0) return false;
    }
    return true;
}

}  // namespace sdt
