// sdt_api.cpp -- the extern "C" boundary of libseidart_hip.so (declared in
// include/seidart_hip.h).  Host staging of the reference's input arrays,
// device setup, the time loop (direct launches or captured hipGraphs of
// several steps for small, launch-bound grids), the blow-up check and the
// read-back of fields, receivers and norms.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "sdt_internal.h"

using namespace sdt;

struct sdt_batch;

struct sdt_handle {
    Device d;
    sdt_params p{};
    std::vector<double> staged[SDT_NUM_ARRAYS];
    std::vector<double> series;       // [SDT_NUM_SRC][time_steps]
    std::vector<int32_t> rec_ijk;     // padded 1-based (i, j, k) per receiver
    std::vector<double> coef;         // update coefficients (sdt_set_coefficients)
    PlaneWaveSpec pw;                 // sdt_set_plane_wave
    bool prepared = false;
    bool failed = false;
    int steps_done = 0;
    std::string err;
    hipGraphExec_t graph[2] = {};      // captured steps, per buffer parity (Device::parity)
    int graph_steps = 0;
    std::vector<hipEvent_t> tev;       // timing events, 3 per timed step
    const sdt_batch *owner = nullptr;  // the batch this handle belongs to (until sdt_batch_destroy)
};

static thread_local std::string g_err;

static int fail(sdt_handle *h, int code, const std::string &msg)
{
    if (h) h->err = msg;
    g_err = msg;
    return code;
}

static int hip_fail(sdt_handle *h, hipError_t e, const char *where)
{
    return fail(h, SDT_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIPCHK(h, call)                                     \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #call); \
    } while (0)

static bool is3d(int physics) { return physics == SDT_ELASTIC3 || physics == SDT_EM3; }
static bool elastic(int physics)
{
    return physics == SDT_ELASTIC2 || physics == SDT_ELASTIC25 || physics == SDT_ELASTIC3;
}
// values per field element: 2 for the complex fields of the 2.5D solver
static int ncomp_of(int physics) { return physics == SDT_ELASTIC25 || physics == SDT_EM25 ? 2 : 1; }

static int nfields_of(int physics, int f)
{
    // fields present per solver
    switch (physics) {
    case SDT_ELASTIC2: return f == SDT_F_VX || f == SDT_F_VZ || f == SDT_F_SXX || f == SDT_F_SZZ || f == SDT_F_SXZ;
    case SDT_ELASTIC25:
    case SDT_ELASTIC3: return f >= 0 && f < 9;
    case SDT_EM2: return f == SDT_F_EX || f == SDT_F_EZ || f == SDT_F_HY;
    case SDT_EM25: return f >= 0 && f < 6;
    case SDT_EM3: return f >= 0 && f < 6;
    }
    return 0;
}

extern "C" {

int32_t sdt_version(void) { return SDT_ABI_VERSION; }

int32_t sdt_device_count(int32_t *count)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (count) *count = (e == hipSuccess) ? n : 0;
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipGetDeviceCount");
    return SDT_OK;
}

const char *sdt_last_error(const sdt_handle *h) { return h ? h->err.c_str() : g_err.c_str(); }

int32_t sdt_create(const sdt_params *p, sdt_handle **out)
{
    if (!p || !out) return fail(nullptr, SDT_ERR_ARG, "null argument");
    *out = nullptr;
    if (p->abi_version != SDT_ABI_VERSION) return fail(nullptr, SDT_ERR_ARG, "abi_version mismatch");
    if (p->physics != SDT_ELASTIC2 && p->physics != SDT_ELASTIC25 && p->physics != SDT_ELASTIC3 &&
        p->physics != SDT_EM2 && p->physics != SDT_EM25 && p->physics != SDT_EM3)
        return fail(nullptr, SDT_ERR_ARG, "unknown physics");
    if (!isfinite(p->ky)) return fail(nullptr, SDT_ERR_ARG, "ky must be finite");
    if (p->precision != SDT_FP64 && p->precision != SDT_FP32) return fail(nullptr, SDT_ERR_ARG, "unknown precision");
    const bool d3 = is3d(p->physics);
    if (p->nx < 5 || p->nz < 5 || (d3 ? p->ny < 5 : p->ny != 1))
        return fail(nullptr, SDT_ERR_ARG, "grid too small (need >= 5 cells per axis; ny = 1 in 2D)");
    if (!(p->dx > 0) || !(p->dz > 0) || (d3 && !(p->dy > 0)) || !(p->dt > 0))
        return fail(nullptr, SDT_ERR_ARG, "dx, dy, dz and dt must be positive");
    if (p->time_steps < 1) return fail(nullptr, SDT_ERR_ARG, "time_steps must be >= 1");
    if (p->cpml < 0) return fail(nullptr, SDT_ERR_ARG, "cpml must be >= 0");
    if (elastic(p->physics) && (p->density_code < 1 || p->density_code > 4))
        return fail(nullptr, SDT_ERR_ARG, "density_code must be 1..4");
    if (p->stencil != SDT_STENCIL_REFERENCE && p->stencil != SDT_STENCIL_CORRECTED)
        return fail(nullptr, SDT_ERR_ARG, "unknown stencil");
    // source cells: padded 1-based, plus the +1 offsets of the elastic stress
    // and velocity sources (cpmlfdtd.f08:519-523, 2208-2219)
    const int is = p->xind + p->cpml, js = (d3 ? p->yind : 1) + (d3 ? p->cpml : 0), ks = p->zind + p->cpml;
    const int off = elastic(p->physics) ? 1 : 0;
    if (is < 1 || ks < 1 || js < 1 || is + off > p->nx || ks + off > p->nz || (d3 && js + off > p->ny))
        return fail(nullptr, SDT_ERR_ARG, "source index outside the padded grid");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev < 1) return fail(nullptr, SDT_ERR_HIP, "no HIP device available");
    if (p->device < 0 || p->device >= ndev) return fail(nullptr, SDT_ERR_ARG, "device ordinal out of range");

    sdt_handle *h = new sdt_handle();
    h->p = *p;
    Device &d = h->d;
    d.physics = p->physics;
    d.precision = p->precision;
    d.esize = p->precision == SDT_FP64 ? 8 : 4;
    d.nx = p->nx; d.ny = p->ny; d.nz = p->nz;
    d.dx = p->dx; d.dy = d3 ? p->dy : 1.0; d.dz = p->dz; d.dt = p->dt;
    d.time_steps = p->time_steps;
    d.isrc = is; d.jsrc = js; d.ksrc = ks;
    d.cpml_width = p->cpml;
    d.density_code = p->density_code;
    d.stencil = p->stencil;
    d.ky = (p->physics == SDT_ELASTIC25 || p->physics == SDT_EM25) ? p->ky : 0.0;
    d.ncomp = ncomp_of(p->physics);
    d.ncell = (int64_t)p->nx * p->ny * p->nz;
    h->series.assign((size_t)SDT_NUM_SRC * p->time_steps, 0.0);
    e = hipSetDevice(p->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking);
    if (e == hipSuccess && p->physics == SDT_ELASTIC3) {
        e = hipStreamCreateWithFlags(&d.side, hipStreamNonBlocking);
        for (int q = 0; q < 2 && e == hipSuccess; ++q) {
            e = hipEventCreateWithFlags(&d.fork[q], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&d.join[q], hipEventDisableTiming);
        }
    }
    if (e != hipSuccess) {
        int rc = hip_fail(nullptr, e, "stream creation");
        delete h;
        return rc;
    }
    *out = h;
    return SDT_OK;
}

static void release(sdt_handle *h)
{
    Device &d = h->d;
    (void)hipSetDevice(h->p.device);
    if (d.stream) (void)hipStreamSynchronize(d.stream);
    for (auto g : h->graph)
        if (g) (void)hipGraphExecDestroy(g);
    auto drop = [](void *p) { if (p) (void)hipFree(p); };
    free_grids(d);   // fields, memory variables, materials, CPML coefficients
    drop(d.tile_flags);
    drop(d.tiles);
    drop(d.mid);   // d.tab is one of the grid allocations freed above
    drop(d.shell_cid);
    drop(d.cpml_cid);   // d.cpml_ctab is a grid allocation
    drop(d.series);
    drop(d.norms);
    drop(d.d_it);
    drop(d.rec_cell);
    drop(d.rec_out);
    drop(d.adj.src_lam);      // the other adjoint arrays are grid allocations
    drop(d.adj.trace_grad);
    drop(d.adj.win_ckpt);
    drop(d.tail.win.v);   // holds the window's coefficients too
    drop(d.snap);
    for (auto ev : h->tev) (void)hipEventDestroy(ev);
    for (int q = 0; q < 2; ++q) {
        if (d.fork[q]) (void)hipEventDestroy(d.fork[q]);
        if (d.join[q]) (void)hipEventDestroy(d.join[q]);
    }
    if (d.side) (void)hipStreamDestroy(d.side);
    if (d.stream) (void)hipStreamDestroy(d.stream);
}

int32_t sdt_destroy(sdt_handle *h)
{
    if (!h) return SDT_OK;
    if (h->owner) return fail(h, SDT_ERR_STATE, "the handle belongs to a batch: destroy the batch first");
    release(h);
    delete h;
    return SDT_OK;
}

int32_t sdt_set_array(sdt_handle *h, int32_t id, const double *host, int64_t count)
{
    if (!h || !host) return fail(h, SDT_ERR_ARG, "null argument");
    if (h->prepared) return fail(h, SDT_ERR_STATE, "inputs are fixed once the first step has run");
    if (id < 0 || id >= SDT_NUM_ARRAYS) return fail(h, SDT_ERR_ARG, "unknown array id");
    const int64_t n = h->d.ncell;
    if (h->d.ncomp == 2 && id >= SDT_IC_X && id <= SDT_IC_Z) {
        // complex initial condition (material_rw2c); a real array sets the real part
        if (count == 2 * n) {
            h->staged[id].assign(host, host + count);
        } else if (count == n) {
            h->staged[id].assign((size_t)(2 * n), 0.0);
            for (int64_t c = 0; c < n; ++c) h->staged[id][(size_t)(2 * c)] = host[c];
        } else {
            return fail(h, SDT_ERR_ARG, "initial-condition length must equal 2*nx*nz (complex) or nx*nz");
        }
        return SDT_OK;
    }
    if (count != n) return fail(h, SDT_ERR_ARG, "array length must equal nx*ny*nz");
    h->staged[id].assign(host, host + count);
    return SDT_OK;
}

int32_t sdt_set_source(sdt_handle *h, int32_t comp, const double *series, int64_t count)
{
    if (!h || !series) return fail(h, SDT_ERR_ARG, "null argument");
    if (h->prepared) return fail(h, SDT_ERR_STATE, "inputs are fixed once the first step has run");
    if (comp < 0 || comp >= SDT_NUM_SRC) return fail(h, SDT_ERR_ARG, "unknown source component");
    if (count != h->d.time_steps) return fail(h, SDT_ERR_ARG, "series length must equal time_steps");
    memcpy(h->series.data() + (size_t)comp * h->d.time_steps, series, (size_t)count * sizeof(double));
    return SDT_OK;
}

int32_t sdt_set_receivers(sdt_handle *h, int32_t n, const int32_t *ijk)
{
    if (!h || (n > 0 && !ijk) || n < 0) return fail(h, SDT_ERR_ARG, "bad receiver list");
    if (h->prepared) return fail(h, SDT_ERR_STATE, "receivers are fixed once the first step has run");
    const Device &d = h->d;
    const bool d3 = is3d(d.physics);
    std::vector<int32_t> cells((size_t)n * 3);
    for (int q = 0; q < n; ++q) {
        const int i = ijk[3 * q] + h->p.cpml;
        const int j = d3 ? ijk[3 * q + 1] + h->p.cpml : 1;
        const int k = ijk[3 * q + 2] + h->p.cpml;
        if (i < 1 || i > d.nx || j < 1 || j > d.ny || k < 1 || k > d.nz)
            return fail(h, SDT_ERR_ARG, "receiver " + std::to_string(q) + " outside the padded grid");
        cells[3 * (size_t)q] = i;
        cells[3 * (size_t)q + 1] = j;
        cells[3 * (size_t)q + 2] = k;
    }
    h->rec_ijk = cells;
    return SDT_OK;
}

static int prepare(sdt_handle *h)
{
    if (h->prepared) return SDT_OK;
    Device &d = h->d;
    HIPCHK(h, hipSetDevice(h->p.device));
    const double *in[SDT_NUM_ARRAYS];
    for (int q = 0; q < SDT_NUM_ARRAYS; ++q) in[q] = h->staged[q].empty() ? nullptr : h->staged[q].data();
    // the reference fills kappa/kappa_half with 1 before reading them
    // (cpmlfdtd.f08:240-247); arrays never set keep that default
    int rc = SDT_OK;
    if (!h->coef.empty() && !d.raw_source)
        return fail(h, SDT_ERR_STATE, "coefficient input needs SDT_OPT_RAW_SOURCE = 1");
    if (d.adj.every > 0 && d.physics != SDT_ELASTIC2 && d.physics != SDT_EM2)
        return fail(h, SDT_ERR_ARG, "checkpointing for the adjoint is available for 2D elastic and 2D EM");
    if (d.adj.every > 0 && h->pw.kind)
        return fail(h, SDT_ERR_ARG, "the adjoint takes point sources, not a plane wave");
    d.coef_in = h->coef.empty() ? nullptr : h->coef.data();
    switch (d.physics) {
    case SDT_ELASTIC2: rc = el2_setup(d, in); break;
    case SDT_ELASTIC25: rc = el25_setup(d, in); break;
    case SDT_ELASTIC3: rc = el3_setup(d, in); break;
    case SDT_EM2: rc = em2_setup(d, in); break;
    case SDT_EM25: rc = em25_setup(d, in); break;
    case SDT_EM3: rc = em3_setup(d, in); break;
    }
    d.coef_in = nullptr;
    if (rc) return fail(h, rc, "device setup failed (out of device memory?)");
    if (h->pw.kind && plane_wave_setup(d, h->pw, in) != SDT_OK)
        return fail(h, SDT_ERR_ARG, (d.physics == SDT_EM2 || d.physics == SDT_EM3)
                                        ? "plane-wave source: the permittivity needs a positive smallest eigenvalue"
                                        : "plane-wave source: the model needs positive density and stiffness");
    if (h->pw.kind && d.tail.win.v) {   // the source window serves point sources only
        (void)hipFree(d.tail.win.v);
        d.tail.win = SrcWindow{};
    }
    if (d.raw_source)
        for (int q = 0; q < d.tail.nterms; ++q) {
            d.tail.terms[q].mul = 1.0;
            d.tail.terms[q].div = 1.0;
        }
    if (d.adj.every > 0) {
        Adjoint &a = d.adj;
        a.nckpt = (d.time_steps + a.every - 1) / a.every + 1;
        const size_t bytes = (size_t)a.nckpt * a.nstate * d.ncell_dev * d.esize;
        void *p = nullptr;
        if (alloc_grid(d, &p, bytes) != SDT_OK) return fail(h, SDT_ERR_HIP, "checkpoint allocation failed");
        a.ckpt = (char *)p;
        if (d.tail.win.v)
            HIPCHK(h, hipMalloc((void **)&a.win_ckpt, (size_t)a.nckpt * src_window_bytes(d.tail.win)));
    }
    // source series, norms, step counter, receivers
    const size_t ns = (size_t)SDT_NUM_SRC * d.time_steps;
    HIPCHK(h, hipMalloc((void **)&d.series, ns * sizeof(double)));
    HIPCHK(h, hipMemcpy(d.series, h->series.data(), ns * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(h, hipMalloc((void **)&d.norms, 2 * (size_t)d.time_steps * sizeof(unsigned long long)));
    HIPCHK(h, hipMemset(d.norms, 0, 2 * (size_t)d.time_steps * sizeof(unsigned long long)));
    HIPCHK(h, hipMalloc((void **)&d.d_it, sizeof(int)));
    const int one = 1;
    HIPCHK(h, hipMemcpy(d.d_it, &one, sizeof(int), hipMemcpyHostToDevice));
    d.nrec = (int)(h->rec_ijk.size() / 3);
    if (d.nrec > 0) {
        std::vector<int64_t> cells((size_t)d.nrec);
        for (int q = 0; q < d.nrec; ++q)
            cells[(size_t)q] = cell_index(d, h->rec_ijk[3 * (size_t)q], h->rec_ijk[3 * (size_t)q + 1],
                                          h->rec_ijk[3 * (size_t)q + 2]);
        HIPCHK(h, hipMalloc((void **)&d.rec_cell, (size_t)d.nrec * sizeof(int64_t)));
        HIPCHK(h, hipMemcpy(d.rec_cell, cells.data(), (size_t)d.nrec * sizeof(int64_t), hipMemcpyHostToDevice));
        const size_t nr = (size_t)d.time_steps * d.nrec * 3 * d.ncomp;
        HIPCHK(h, hipMalloc((void **)&d.rec_out, nr * sizeof(double)));
        HIPCHK(h, hipMemset(d.rec_out, 0, nr * sizeof(double)));
    }
    TailArgs &t = d.tail;
    t.d_it = d.d_it;
    t.series = d.series;
    t.time_steps = d.time_steps;
    t.vnorm = d.norms;
    t.snorm = d.norms + d.time_steps;
    t.nrec = d.nrec;
    t.rec_cell = d.rec_cell;
    t.rec_out = d.rec_out;
    HIPCHK(h, hipDeviceSynchronize());
    for (auto &s : h->staged) { std::vector<double>().swap(s); }
    std::vector<double>().swap(h->coef);
    h->prepared = true;
    return SDT_OK;
}

static void enqueue_step(sdt_handle *h, bool timed)
{
    Device &d = h->d;
    switch (d.physics) {
    case SDT_ELASTIC2: el2_step(d, timed); break;
    case SDT_ELASTIC25: el25_step(d, timed); break;
    case SDT_ELASTIC3: el3_step(d, timed); break;
    case SDT_EM2: em2_step(d, timed); break;
    case SDT_EM25: em25_step(d, timed); break;
    case SDT_EM3: em3_step(d, timed); break;
    }
}

// Capture kGraphSteps steps once and replay them: no per-kernel launch gaps
// (small grids are launch-bound; large ones still lose a few microseconds per
// launch).  The step index lives in device memory, so one graph serves every
// replay.
static constexpr int kGraphSteps = 32;

static void drop_graphs(sdt_handle *h)
{
    for (auto &g : h->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

static int ensure_graph(sdt_handle *h)
{
    Device &d = h->d;
    if (h->graph[d.parity]) return SDT_OK;
    const int par = d.parity;
    hipGraph_t g = nullptr;
    HIPCHK(h, hipStreamBeginCapture(d.stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < kGraphSteps; ++s) enqueue_step(h, false);
    HIPCHK(h, hipStreamEndCapture(d.stream, &g));
    hipError_t e = hipGraphInstantiate(&h->graph[par], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(h, e, "hipGraphInstantiate");
    h->graph_steps = kGraphSteps;
    return SDT_OK;
}

// checkpoint slot k <-> the state arrays (fields + memory variables), on
// stream s (the handle's, or its batch's)
static void checkpoint_copy(Device &d, int k, bool save, hipStream_t s)
{
    const Adjoint &a = d.adj;
    const size_t bytes = (size_t)d.ncell_dev * d.esize;
    for (int q = 0; q < a.nstate; ++q) {
        char *slot = a.ckpt + ((size_t)k * a.nstate + q) * bytes;
        if (save) (void)hipMemcpyAsync(slot, a.state[q], bytes, hipMemcpyDeviceToDevice, s);
        else (void)hipMemcpyAsync(a.state[q], slot, bytes, hipMemcpyDeviceToDevice, s);
    }
    if (a.win_ckpt) {   // the FP64 source window is part of the state
        const size_t wb = src_window_bytes(d.tail.win);
        char *slot = (char *)a.win_ckpt + (size_t)k * wb;
        if (save) (void)hipMemcpyAsync(slot, d.tail.win.v, wb, hipMemcpyDeviceToDevice, s);
        else (void)hipMemcpyAsync(d.tail.win.v, slot, wb, hipMemcpyDeviceToDevice, s);
    }
}

// Check the velocity norms of steps [first, first+n) (1-based first).
static int check_norms(sdt_handle *h, int first, int n, int32_t *failed_step)
{
    Device &d = h->d;
    std::vector<unsigned long long> bits((size_t)n);
    HIPCHK(h, hipMemcpyAsync(bits.data(), d.norms + (first - 1), (size_t)n * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, d.stream));
    HIPCHK(h, hipStreamSynchronize(d.stream));
    for (int s = 0; s < n; ++s) {
        double v;
        memcpy(&v, &bits[(size_t)s], sizeof(double));
        if (v > kStability || v != v) {
            if (failed_step) *failed_step = first + s;
            h->failed = true;
            h->steps_done = first + s;
            return fail(h, SDT_ERR_UNSTABLE, "code became unstable and blew up at step " + std::to_string(first + s));
        }
    }
    return SDT_OK;
}

int32_t sdt_run(sdt_handle *h, int32_t nsteps, int32_t *failed_step)
{
    if (!h) return fail(nullptr, SDT_ERR_ARG, "null handle");
    if (failed_step) *failed_step = 0;
    if (h->owner) return fail(h, SDT_ERR_STATE, "the handle belongs to a batch: advance it with sdt_batch_run");
    if (h->failed) return fail(h, SDT_ERR_STATE, "the run has already blown up");
    if (nsteps < 0) return fail(h, SDT_ERR_ARG, "nsteps must be >= 0");
    if (h->steps_done + nsteps > h->d.time_steps)
        return fail(h, SDT_ERR_ARG, "run would exceed time_steps");
    int rc = prepare(h);
    if (rc) return rc;
    Device &d = h->d;
    HIPCHK(h, hipSetDevice(h->p.device));
    const bool use_graph = d.graphs && !d.timing;
    const int chunk = 256;   // steps between blow-up checks
    int left = nsteps;
    const int every = d.adj.every;
    while (left > 0) {
        int n = left < chunk ? left : chunk;
        if (every > 0) {   // segments end at multiples of `every`; save each segment's first state
            const int into = h->steps_done % every;
            if (into == 0) checkpoint_copy(d, h->steps_done / every, true, d.stream);
            n = n < every - into ? n : every - into;
        }
        const int first = h->steps_done + 1;
        if (d.timing) {
            if ((int)h->tev.size() < 3 * n) {
                size_t old = h->tev.size();
                h->tev.resize((size_t)3 * n);
                for (size_t q = old; q < h->tev.size(); ++q) HIPCHK(h, hipEventCreate(&h->tev[q]));
            }
            for (int s = 0; s < n; ++s) {
                d.ev[0] = h->tev[3 * s];
                d.ev[1] = h->tev[3 * s + 1];
                d.ev[2] = h->tev[3 * s + 2];
                enqueue_step(h, true);
            }
            HIPCHK(h, hipStreamSynchronize(d.stream));
            for (int s = 0; s < n; ++s) {
                float a = 0, b = 0;
                HIPCHK(h, hipEventElapsedTime(&a, h->tev[3 * s], h->tev[3 * s + 1]));
                HIPCHK(h, hipEventElapsedTime(&b, h->tev[3 * s + 1], h->tev[3 * s + 2]));
                d.ms_stress += a;
                d.ms_vel += b;
                d.timed_launches += 1;
            }
        } else {
            int s = 0;
            if (use_graph && n >= kGraphSteps) {
                rc = ensure_graph(h);
                if (rc) return rc;
                for (; s + kGraphSteps <= n; s += kGraphSteps) HIPCHK(h, hipGraphLaunch(h->graph[d.parity], d.stream));
            }
            for (; s < n; ++s) enqueue_step(h, false);
        }
        HIPCHK(h, hipGetLastError());
        rc = check_norms(h, first, n, failed_step);
        if (rc) return rc;
        h->steps_done += n;
        left -= n;
        if (every > 0 && h->steps_done == d.time_steps) checkpoint_copy(d, d.adj.nckpt - 1, true, d.stream);
    }
    return SDT_OK;
}

int32_t sdt_steps_done(const sdt_handle *h, int32_t *steps)
{
    if (!h || !steps) return fail(nullptr, SDT_ERR_ARG, "null argument");
    *steps = h->steps_done;
    return SDT_OK;
}

int32_t sdt_get_field(sdt_handle *h, int32_t f, double *out, int64_t count)
{
    if (!h || !out) return fail(h, SDT_ERR_ARG, "null argument");
    if (f < 0 || f >= SDT_NUM_FIELDS || !nfields_of(h->d.physics, f)) return fail(h, SDT_ERR_ARG, "field not present in this solver");
    if (count != h->d.ncell * h->d.ncomp)
        return fail(h, SDT_ERR_ARG, h->d.ncomp == 2 ? "count must equal 2*nx*nz (complex)" : "count must equal nx*ny*nz");
    int rc = prepare(h);
    if (rc) return rc;
    Device &d = h->d;
    HIPCHK(h, hipSetDevice(h->p.device));
    HIPCHK(h, hipStreamSynchronize(d.stream));
    if (download_grid(d, d.field[f], out, d.ncomp) != SDT_OK) return fail(h, SDT_ERR_HIP, "field read-back failed");
    // lazily applied Dirichlet faces of H (em2.hip / em3.hip)
    const bool hfield = (d.physics == SDT_EM2 && f == SDT_F_HY) || (d.physics == SDT_EM3 && f >= SDT_F_HX);
    if (hfield && h->steps_done > 0) {
        for (int k = 1; k <= d.nz; ++k)
            for (int j = 1; j <= d.ny; ++j)
                for (int i = 1; i <= d.nx; ++i) {
                    const bool face = i == 1 || i == d.nx || k == 1 || k == d.nz ||
                                      (d.physics == SDT_EM3 && (j == 1 || j == d.ny));
                    if (face) out[(int64_t)(i - 1) + (int64_t)d.nx * ((int64_t)(j - 1) + (int64_t)d.ny * (k - 1))] = 0.0;
                }
    }
    return SDT_OK;
}

int32_t sdt_set_field(sdt_handle *h, int32_t f, const double *host, int64_t count)
{
    if (!h || !host) return fail(h, SDT_ERR_ARG, "null argument");
    if (f < 0 || f >= SDT_NUM_FIELDS || !nfields_of(h->d.physics, f)) return fail(h, SDT_ERR_ARG, "field not present in this solver");
    if (count != h->d.ncell * h->d.ncomp)
        return fail(h, SDT_ERR_ARG, h->d.ncomp == 2 ? "count must equal 2*nx*nz (complex)" : "count must equal nx*ny*nz");
    int rc = prepare(h);
    if (rc) return rc;
    Device &d = h->d;
    HIPCHK(h, hipSetDevice(h->p.device));
    HIPCHK(h, hipStreamSynchronize(d.stream));
    if (store_grid(d, d.field[f], host, d.ncomp) != SDT_OK) return fail(h, SDT_ERR_HIP, "field upload failed");
    if (src_window_load(d, f, host) != SDT_OK) return fail(h, SDT_ERR_HIP, "field upload failed");
    return SDT_OK;
}

// H fields whose Dirichlet faces the EM kernels apply lazily (em2.hip, em3.hip)
static bool lazy_h_faces(const sdt_handle *h, int f)
{
    const Device &d = h->d;
    const bool hfield = (d.physics == SDT_EM2 && f == SDT_F_HY) || (d.physics == SDT_EM3 && f >= SDT_F_HX);
    return hfield && h->steps_done > 0;
}

static int snapshot_check(sdt_handle *h, int32_t f, int32_t single, const void *out, int64_t count)
{
    if (!h || !out) return fail(h, SDT_ERR_ARG, "null argument");
    if (f < 0 || f >= SDT_NUM_FIELDS || !nfields_of(h->d.physics, f))
        return fail(h, SDT_ERR_ARG, "field not present in this solver");
    if (h->d.ncomp != 1) return fail(h, SDT_ERR_ARG, "snapshots are written for the real-valued solvers");
    if (single != 0 && single != 1) return fail(h, SDT_ERR_ARG, "single must be 0 (FP64) or 1 (FP32)");
    if (count != h->d.ncell) return fail(h, SDT_ERR_ARG, "count must equal nx*ny*nz");
    return SDT_OK;
}

int32_t sdt_snapshot(sdt_handle *h, int32_t f, int32_t single, void *out, int64_t count)
{
    int rc = snapshot_check(h, f, single, out, count);
    if (rc) return rc;
    rc = prepare(h);
    if (rc) return rc;
    Device &d = h->d;
    HIPCHK(h, hipSetDevice(h->p.device));
    if (!d.snap) HIPCHK(h, hipMalloc(&d.snap, (size_t)d.ncell * sizeof(double)));
    launch_pack(d, d.field[f], d.snap, single, 0, d.nz, lazy_h_faces(h, f), d.stream);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(out, d.snap, (size_t)count * (single ? 4 : 8), hipMemcpyDeviceToHost, d.stream));
    return SDT_OK;
}

// ---- allocation-free snapshots ----------------------------------------------
// Device scratch of sdt_snapshot_noalloc: one per GPU, held until the process
// ends.  Host code never uses the symbol itself (in host code it is only a
// stand-in, not a device address): the device address is looked up once per
// device with that device current and kept in g_scratch_ptr.
constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];
constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

static int scratch_for(sdt_handle *h, void **ptr)
{
    const int dev = h->p.device;
    if (dev < 0 || dev >= kMaxDevices) return fail(h, SDT_ERR_ARG, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        HIPCHK(h, hipSetDevice(dev));
        void *p = nullptr;
        HIPCHK(h, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return SDT_OK;
}

int32_t sdt_snapshot_noalloc(sdt_handle *h, int32_t f, int32_t single, void *out, int64_t count)
{
    int rc = snapshot_check(h, f, single, out, count);
    if (rc) return rc;
    if (!h->prepared)
        return fail(h, SDT_ERR_STATE, "run the handle first (sdt_run with 0 steps sets it up): "
                                      "this entry point allocates nothing");
    Device &d = h->d;
    const size_t esz = single ? 4 : 8;
    const size_t plane = (size_t)d.nx * d.ny * esz;
    if (plane > kScratchBytes)
        return fail(h, SDT_ERR_ARG, "one x-y plane (" + std::to_string(plane) + " bytes) exceeds the " +
                                        std::to_string(kScratchBytes >> 20) + " MiB snapshot scratch");
    void *ptr = nullptr;
    rc = scratch_for(h, &ptr);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->p.device));
    // planes in chunks that fit the scratch; on one stream the next chunk's
    // pack waits for the previous chunk's copy out of the scratch
    const int per = (int)(kScratchBytes / plane);
    char *dst = (char *)out;
    for (int k0 = 0; k0 < d.nz; k0 += per) {
        const int nk = d.nz - k0 < per ? d.nz - k0 : per;
        launch_pack(d, d.field[f], ptr, single, k0, nk, lazy_h_faces(h, f), d.stream);
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipMemcpyAsync(dst + (size_t)k0 * plane, ptr, (size_t)nk * plane, hipMemcpyDeviceToHost,
                                 d.stream));
    }
    HIPCHK(h, hipStreamSynchronize(d.stream));
    return SDT_OK;
}

int32_t sdt_host_alloc(int64_t bytes, void **out)
{
    if (!out || bytes <= 0) return fail(nullptr, SDT_ERR_ARG, "bad pinned allocation request");
    *out = nullptr;
    hipError_t e = hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipHostMalloc");
    return SDT_OK;
}

int32_t sdt_host_free(void *p)
{
    if (!p) return SDT_OK;
    hipError_t e = hipHostFree(p);
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipHostFree");
    return SDT_OK;
}

int32_t sdt_get_receivers(sdt_handle *h, double *out, int64_t count)
{
    if (!h || (!out && count)) return fail(h, SDT_ERR_ARG, "null argument");
    const int64_t need = (int64_t)h->steps_done * (int64_t)(h->rec_ijk.size() / 3) * 3 * h->d.ncomp;
    if (count != need)
        return fail(h, SDT_ERR_ARG, h->d.ncomp == 2 ? "count must equal steps_done * n_receivers * 6 (complex)"
                                                    : "count must equal steps_done * n_receivers * 3");
    if (need == 0) return SDT_OK;
    HIPCHK(h, hipSetDevice(h->p.device));
    HIPCHK(h, hipStreamSynchronize(h->d.stream));
    HIPCHK(h, hipMemcpy(out, h->d.rec_out, (size_t)need * sizeof(double), hipMemcpyDeviceToHost));
    return SDT_OK;
}

int32_t sdt_get_norms(sdt_handle *h, double *vel, double *stress, int64_t count)
{
    if (!h) return fail(nullptr, SDT_ERR_ARG, "null handle");
    if (count != h->steps_done) return fail(h, SDT_ERR_ARG, "count must equal steps_done");
    if (count == 0) return SDT_OK;
    HIPCHK(h, hipSetDevice(h->p.device));
    HIPCHK(h, hipStreamSynchronize(h->d.stream));
    std::vector<unsigned long long> bits((size_t)count);
    double *dst[2] = {vel, stress};
    for (int q = 0; q < 2; ++q) {
        if (!dst[q]) continue;
        HIPCHK(h, hipMemcpy(bits.data(), h->d.norms + (size_t)q * h->d.time_steps, (size_t)count * 8, hipMemcpyDeviceToHost));
        memcpy(dst[q], bits.data(), (size_t)count * 8);
    }
    return SDT_OK;
}

int32_t sdt_set_timing(sdt_handle *h, int32_t enable)
{
    if (!h) return fail(nullptr, SDT_ERR_ARG, "null handle");
    h->d.timing = enable != 0;
    h->d.ms_stress = h->d.ms_vel = 0;
    h->d.timed_launches = 0;
    return SDT_OK;
}

int32_t sdt_get_kernel_times(sdt_handle *h, double *ms_stress, double *ms_velocity, int64_t *launches)
{
    if (!h) return fail(nullptr, SDT_ERR_ARG, "null handle");
    if (ms_stress) *ms_stress = h->d.ms_stress;
    if (ms_velocity) *ms_velocity = h->d.ms_vel;
    if (launches) *launches = h->d.timed_launches;
    return SDT_OK;
}

int32_t sdt_set_option(sdt_handle *h, int32_t key, int64_t value)
{
    if (!h) return fail(nullptr, SDT_ERR_ARG, "null handle");
    switch (key) {
    case SDT_OPT_ZCHUNK:
        if (h->prepared) return fail(h, SDT_ERR_STATE, "z chunk is fixed once the first step has run");
        if (value < 0 || value > 1 << 20) return fail(h, SDT_ERR_ARG, "z chunk out of range");
        h->d.zchunk = (int)value;
        return SDT_OK;
    case SDT_OPT_TILE_ROWS:
        if (value != 0 && value != 4 && value != 8 && value != 16)
            return fail(h, SDT_ERR_ARG, "tile rows must be 0 (auto), 4, 8 or 16");
        h->d.tile_rows = (int)value;
        return SDT_OK;
    case SDT_OPT_GRAPH:
        h->d.graphs = value != 0;
        return SDT_OK;
    case SDT_OPT_CONCURRENT:
        h->d.concurrent = value != 0;
        drop_graphs(h);   // captured steps follow the old setting: capture again
        return SDT_OK;
    case SDT_OPT_XCD_REMAP:
        if (value < 0 || value > 3) return fail(h, SDT_ERR_ARG, "XCD remap option must be 0..3");
        h->d.xcd_remap = (int)value;
        drop_graphs(h);
        return SDT_OK;
    case SDT_OPT_MATERIAL_TABLE:
        if (h->prepared) return fail(h, SDT_ERR_STATE, "material layout is fixed once the first step has run");
        if (value != 0 && value != 1) return fail(h, SDT_ERR_ARG, "material table option must be 0 or 1");
        h->d.mat_table_opt = (int)value;
        return SDT_OK;
    case SDT_OPT_FUSED:
        if (h->prepared) return fail(h, SDT_ERR_STATE, "the step's kernels are fixed once the first step has run");
        if (value < -1 || value > 1) return fail(h, SDT_ERR_ARG, "fused option must be -1, 0 or 1");
        h->d.fused_opt = (int)value;
        return SDT_OK;
    case SDT_OPT_ISOTROPIC:
        if (h->prepared) return fail(h, SDT_ERR_STATE, "material layout is fixed once the first step has run");
        if (value != 0 && value != 1) return fail(h, SDT_ERR_ARG, "isotropic option must be 0 or 1");
        h->d.iso_opt = (int)value;
        return SDT_OK;
    case SDT_OPT_STAGGER:
        if (h->prepared) return fail(h, SDT_ERR_STATE, "allocation layout is fixed once the first step has run");
        h->d.stagger = value != 0;
        return SDT_OK;
    case SDT_OPT_CHECKPOINT:
        if (h->prepared) return fail(h, SDT_ERR_STATE, "checkpointing is fixed once the first step has run");
        if (value < 0 || value > h->d.time_steps) return fail(h, SDT_ERR_ARG, "checkpoint interval out of range");
        h->d.adj.every = (int)value;
        return SDT_OK;
    case SDT_OPT_RAW_SOURCE:
        if (h->prepared) return fail(h, SDT_ERR_STATE, "source scaling is fixed once the first step has run");
        h->d.raw_source = value != 0;
        return SDT_OK;
    case SDT_OPT_SOURCE_WINDOW:
        if (h->prepared) return fail(h, SDT_ERR_STATE, "the source window is fixed once the first step has run");
        if (value != 0 && value != 1) return fail(h, SDT_ERR_ARG, "source window option must be 0 or 1");
        h->d.src_window_opt = value != 0;
        return SDT_OK;
    }
    return fail(h, SDT_ERR_ARG, "unknown option");
}

int32_t sdt_get_info(sdt_handle *h, int32_t key, int64_t *value)
{
    if (!h || !value) return fail(h, SDT_ERR_ARG, "null argument");
    if (!h->prepared) return fail(h, SDT_ERR_STATE, "layout is chosen at the first step");
    switch (key) {
    case SDT_INFO_MATERIALS: *value = h->d.mat_table ? h->d.nmaterials : 0; return SDT_OK;
    case SDT_INFO_SHELL_TILES: *value = h->d.ntiles; return SDT_OK;
    case SDT_INFO_DEVICE_BYTES: *value = (int64_t)h->d.grid_bytes; return SDT_OK;
    case SDT_INFO_COEFFICIENTS: *value = h->d.adj.ncoef; return SDT_OK;
    case SDT_INFO_ISOTROPIC: *value = h->d.mat_iso ? 1 : 0; return SDT_OK;
    case SDT_INFO_FUSED: *value = h->d.fused ? 1 : 0; return SDT_OK;
    case SDT_INFO_CPML_TUPLES: *value = h->d.physics == SDT_ELASTIC3 ? h->d.shell_ncoef : h->d.cpml_ncoef; return SDT_OK;
    }
    return fail(h, SDT_ERR_ARG, "unknown info key");
}

// ---- adjoint-state gradients ---------------------------------------------------
static int ncoef_of(int physics)
{
    return physics == SDT_EM2 ? 7 : physics == SDT_ELASTIC2 ? 10 : 0;
}

int32_t sdt_set_plane_wave(sdt_handle *h, int32_t kind, double angle_xz, double angle_xy, double angle_yz, double f0,
                           double amplitude, int32_t wavelet)
{
    if (!h) return fail(nullptr, SDT_ERR_ARG, "null handle");
    if (h->prepared) return fail(h, SDT_ERR_STATE, "inputs are fixed once the first step has run");
    const int ph = h->d.physics;
    if (ph == SDT_EM25)   // cpmlfdtd.f08:2937-2996 sets up no injection for it
        return fail(h, SDT_ERR_ARG, "the 2.5D electromagnetic solver has no plane-wave source");
    if (kind != SDT_PW_S && kind != SDT_PW_P) return fail(h, SDT_ERR_ARG, "plane-wave kind must be SDT_PW_S or SDT_PW_P");
    if ((ph == SDT_EM2 || ph == SDT_EM3) && kind != SDT_PW_S)   // source type 'pw' only
        return fail(h, SDT_ERR_ARG, "electromagnetic plane waves are transverse: kind SDT_PW_S");
    if (!(f0 > 0) || !isfinite(f0) || !isfinite(amplitude) || !isfinite(angle_xz) || !isfinite(angle_xy) ||
        !isfinite(angle_yz))
        return fail(h, SDT_ERR_ARG, "plane-wave parameters must be finite, f0 > 0");
    h->pw.kind = kind;
    h->pw.angle_xz = angle_xz;
    h->pw.angle_xy = angle_xy;
    h->pw.angle_yz = angle_yz;
    h->pw.f0 = f0;
    h->pw.amplitude = amplitude;
    h->pw.wavelet = wavelet >= 0 && wavelet <= 2 ? wavelet : 0;
    return SDT_OK;
}

int32_t sdt_set_coefficients(sdt_handle *h, const double *coef, int64_t count)
{
    if (!h || !coef) return fail(h, SDT_ERR_ARG, "null argument");
    if (h->prepared) return fail(h, SDT_ERR_STATE, "inputs are fixed once the first step has run");
    const int nc = ncoef_of(h->d.physics);
    if (!nc) return fail(h, SDT_ERR_ARG, "coefficient input is available for 2D elastic and 2D EM");
    if (count != (int64_t)nc * h->d.ncell) return fail(h, SDT_ERR_ARG, "count must equal ncoef * nx * nz");
    h->coef.assign(coef, coef + count);
    return SDT_OK;
}

// adjoint arrays: segment stores, gradient, source and trace buffers, the
// solver's adjoint fields
static int adjoint_alloc(sdt_handle *h)
{
    Device &d = h->d;
    Adjoint &a = d.adj;
    if (a.grad) return SDT_OK;
    const size_t cell = (size_t)d.ncell_dev * d.esize;
    void *p = nullptr;
    if (alloc_grid(d, &p, (size_t)a.every * a.nfwd * cell) != SDT_OK) return SDT_ERR_HIP;
    a.seg_fwd = p;
    if (alloc_grid(d, &p, (size_t)a.every * a.nadj * cell) != SDT_OK) return SDT_ERR_HIP;
    a.seg_adj = p;
    if (alloc_grid(d, &p, (size_t)a.ngrad * d.ncell_dev * sizeof(double)) != SDT_OK) return SDT_ERR_HIP;
    a.grad = (double *)p;
    if (hipMalloc((void **)&a.src_lam, (size_t)d.time_steps * kAdjSrcSlots * sizeof(double)) != hipSuccess)
        return SDT_ERR_HIP;
    if (hipMalloc((void **)&a.trace_grad, ((size_t)d.time_steps * (d.nrec > 0 ? d.nrec : 1) * 3) * sizeof(double)) !=
        hipSuccess)
        return SDT_ERR_HIP;
    return d.physics == SDT_EM2 ? em2_adj_setup(d) : el2_adj_setup(d);
}

// zero the adjoint state and upload the trace gradient
static int adjoint_reset(sdt_handle *h, const double *trace_grad)
{
    Device &d = h->d;
    Adjoint &a = d.adj;
    const size_t cell = (size_t)d.ncell_dev * d.esize;
    for (int q = 0; q < a.nlam; ++q) HIPCHK(h, hipMemsetAsync(a.lam[q], 0, cell, d.stream));
    HIPCHK(h, hipMemsetAsync(a.grad, 0, (size_t)a.ngrad * d.ncell_dev * sizeof(double), d.stream));
    HIPCHK(h, hipMemsetAsync(a.src_lam, 0, (size_t)d.time_steps * kAdjSrcSlots * sizeof(double), d.stream));
    if (d.nrec > 0)
        HIPCHK(h, hipMemcpyAsync(a.trace_grad, trace_grad, (size_t)d.time_steps * d.nrec * 3 * sizeof(double),
                                 hipMemcpyHostToDevice, d.stream));
    HIPCHK(h, hipStreamSynchronize(d.stream));
    return SDT_OK;
}

static void adjoint_step(Device &d, int t)
{
    if (d.physics == SDT_EM2) em2_adj_step(d, t);
    else el2_adj_step(d, t);
}

static void adjoint_correlate(Device &d, int first, int n)
{
    if (d.physics == SDT_EM2) em2_adj_correlate(d, first, n);
    else el2_adj_correlate(d, first, n);
}

// gradient and source gradient back to the host
static int adjoint_download(sdt_handle *h, double *coef_grad, double *src_grad)
{
    Device &d = h->d;
    Adjoint &a = d.adj;
    HIPCHK(h, hipStreamSynchronize(d.stream));
    for (int q = 0; q < a.ncoef; ++q)
        if (download_grid_f64(d, a.grad + (size_t)q * d.ncell_dev, coef_grad + (size_t)q * d.ncell) != SDT_OK)
            return fail(h, SDT_ERR_HIP, "gradient read-back failed");
    if (src_grad) {
        std::vector<double> lam((size_t)d.time_steps * kAdjSrcSlots);
        HIPCHK(h, hipMemcpy(lam.data(), a.src_lam, lam.size() * sizeof(double), hipMemcpyDeviceToHost));
        memset(src_grad, 0, (size_t)SDT_NUM_SRC * d.time_steps * sizeof(double));
        // term k of the tail records its field's adjoint in slot k; the series
        // enters as (v * mul) / div
        for (int k = 0; k < d.tail.nterms && k < kAdjSrcSlots; ++k) {
            const SrcTerm &s = d.tail.terms[k];
            if (s.cell < 0 || s.on_face) continue;
            const double scale = s.mul / s.div;
            for (int t = 0; t < d.time_steps; ++t)
                src_grad[(size_t)s.component * d.time_steps + t] += lam[(size_t)t * kAdjSrcSlots + k] * scale;
        }
    }
    return SDT_OK;
}

int32_t sdt_adjoint(sdt_handle *h, const double *trace_grad, int64_t tcount, double *coef_grad, int64_t gcount,
                    double *src_grad, int64_t scount)
{
    if (!h || !coef_grad || (!trace_grad && tcount)) return fail(h, SDT_ERR_ARG, "null argument");
    if (h->owner) return fail(h, SDT_ERR_STATE, "the handle belongs to a batch: use sdt_batch_adjoint");
    Device &d = h->d;
    Adjoint &a = d.adj;
    if (d.physics != SDT_EM2 && d.physics != SDT_ELASTIC2)
        return fail(h, SDT_ERR_ARG, "the adjoint is available for 2D elastic and 2D EM");
    if (!h->prepared || a.every <= 0 || h->steps_done != d.time_steps)
        return fail(h, SDT_ERR_STATE, "run all time_steps with SDT_OPT_CHECKPOINT set first");
    if (h->failed) return fail(h, SDT_ERR_STATE, "the run has blown up");
    if (tcount != (int64_t)d.time_steps * d.nrec * 3) return fail(h, SDT_ERR_ARG, "tcount must equal time_steps * n_receivers * 3");
    if (gcount != (int64_t)a.ncoef * d.ncell) return fail(h, SDT_ERR_ARG, "gcount must equal ncoef * nx * nz");
    if (src_grad && scount != (int64_t)SDT_NUM_SRC * d.time_steps)
        return fail(h, SDT_ERR_ARG, "scount must equal SDT_NUM_SRC * time_steps");
    HIPCHK(h, hipSetDevice(h->p.device));
    if (adjoint_alloc(h) != SDT_OK) return fail(h, SDT_ERR_HIP, "adjoint allocation failed (out of device memory?)");
    int rc = adjoint_reset(h, trace_grad);
    if (rc) return rc;
    const int nseg = a.nckpt - 1;
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int first = seg * a.every;
        const int n = d.time_steps - first < a.every ? d.time_steps - first : a.every;
        checkpoint_copy(d, seg, false, d.stream);
        set_step_counter(d.stream, d.d_it, first + 1);
        a.store = true;
        a.store_first = first;
        for (int s = 0; s < n; ++s) enqueue_step(h, false);
        a.store = false;
        for (int t = first + n - 1; t >= first; --t) adjoint_step(d, t);
        adjoint_correlate(d, first, n);
        HIPCHK(h, hipGetLastError());
    }
    if (d.physics == SDT_ELASTIC2) el2_adj_finish(d);
    checkpoint_copy(d, a.nckpt - 1, false, d.stream);   // back to the forward run's final state
    set_step_counter(d.stream, d.d_it, d.time_steps + 1);
    HIPCHK(h, hipGetLastError());
    return adjoint_download(h, coef_grad, src_grad);
}

// ---- batched shots ----------------------------------------------------------
struct sdt_batch {
    std::vector<sdt_handle *> shots;
    Batch b;
    hipGraphExec_t graph = nullptr;
    hipEvent_t ev[2] = {};
    std::string err;
};

static int bfail(sdt_batch *b, int code, const std::string &msg)
{
    if (b) b->err = msg;
    g_err = msg;
    return code;
}

int32_t sdt_batch_create(sdt_handle *const *handles, int32_t n, sdt_batch **out)
{
    if (!handles || !out || n < 1) return fail(nullptr, SDT_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    // a handle listed twice would have its tiles updated twice per launch
    // (racing on its fields) and its sources injected twice per step; the
    // check compares pointers only, before any handle is dereferenced
    for (int q = 0; q < n; ++q) {
        if (!handles[q]) return fail(nullptr, SDT_ERR_ARG, "null handle in batch");
        for (int r = 0; r < q; ++r)
            if (handles[r] == handles[q])
                return fail(nullptr, SDT_ERR_ARG,
                            "handle " + std::to_string(q) + " repeats handle " + std::to_string(r) + " in the batch");
    }
    const int phys = handles[0] ? handles[0]->p.physics : -1;
    if (phys != SDT_ELASTIC2 && phys != SDT_EM2)
        return fail(nullptr, SDT_ERR_ARG, "batches hold 2D elastic or 2D EM handles");
    for (int q = 0; q < n; ++q) {
        sdt_handle *h = handles[q];
        if (!h) return fail(nullptr, SDT_ERR_ARG, "null handle in batch");
        if (h->p.physics != phys || h->p.precision != handles[0]->p.precision ||
            h->p.device != handles[0]->p.device)
            return fail(nullptr, SDT_ERR_ARG, "batch handles must share solver kind, precision and device");
        if (h->failed) return fail(nullptr, SDT_ERR_STATE, "a batch handle has already blown up");
        if (h->owner) return fail(nullptr, SDT_ERR_STATE, "a handle already belongs to another batch");
        if (h->pw.kind) return fail(nullptr, SDT_ERR_ARG, "batched shots take point sources, not a plane wave");
        int rc = prepare(h);
        if (rc) return rc;
    }
    sdt_batch *b = new sdt_batch();
    b->shots.assign(handles, handles + n);
    std::vector<Device *> devs((size_t)n);
    for (int q = 0; q < n; ++q) devs[(size_t)q] = &handles[q]->d;
    hipError_t e = hipSetDevice(handles[0]->p.device);
    int rc = e == hipSuccess ? batch_setup(b->b, devs.data(), n) : SDT_ERR_HIP;
    if (rc == SDT_OK) {
        for (int q = 0; q < 2 && rc == SDT_OK; ++q)
            if (hipEventCreate(&b->ev[q]) != hipSuccess) rc = SDT_ERR_HIP;
    }
    if (rc != SDT_OK) {
        batch_free(b->b);
        delete b;
        return fail(nullptr, rc, "batch setup failed");
    }
    for (sdt_handle *h : b->shots) h->owner = b;
    *out = b;
    return SDT_OK;
}

// advance every unfinished shot of the batch `todo` steps on the batch's
// stream: tile-shape tuning on the first steps (each candidate runs one
// untimed and kTune timed real steps; every shape gives the same results),
// then replays of a captured kSteps-step graph, then single steps
static int batch_advance(sdt_batch *b, int todo)
{
    Batch &B = b->b;
    hipError_t e;
    int done = 0;
    constexpr int kTune = 4;
    if (B.choice < 0 && todo >= kNumBatchShapes * (kTune + 1)) {
        float best = 0;
        for (int w = 0; w < kNumBatchShapes; ++w) {
            batch_step(B, w);
            (void)hipEventRecord(b->ev[0], B.stream);
            for (int s = 0; s < kTune; ++s) batch_step(B, w);
            (void)hipEventRecord(b->ev[1], B.stream);
            if ((e = hipEventSynchronize(b->ev[1])) != hipSuccess) return bfail(b, SDT_ERR_HIP, hipGetErrorString(e));
            float ms = 0;
            (void)hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
            if (B.choice < 0 || ms < best) { best = ms; B.choice = w; }
        }
        done += kNumBatchShapes * (kTune + 1);
    }
    const int w = B.choice < 0 ? 1 : B.choice;
    constexpr int kSteps = 32;
    if (todo - done >= kSteps && !b->graph) {
        hipGraph_t g = nullptr;
        if ((e = hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return bfail(b, SDT_ERR_HIP, hipGetErrorString(e));
        for (int s = 0; s < kSteps; ++s) batch_step(B, w);
        if ((e = hipStreamEndCapture(B.stream, &g)) != hipSuccess) return bfail(b, SDT_ERR_HIP, hipGetErrorString(e));
        e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return bfail(b, SDT_ERR_HIP, hipGetErrorString(e));
    }
    for (; b->graph && done + kSteps <= todo; done += kSteps)
        if ((e = hipGraphLaunch(b->graph, B.stream)) != hipSuccess) return bfail(b, SDT_ERR_HIP, hipGetErrorString(e));
    for (; done < todo; ++done) batch_step(B, w);
    return SDT_OK;
}

int32_t sdt_batch_run(sdt_batch *b, int32_t nsteps, int32_t *failed_shot, int32_t *failed_step)
{
    if (!b) return fail(nullptr, SDT_ERR_ARG, "null batch");
    if (failed_shot) *failed_shot = -1;
    if (failed_step) *failed_step = 0;
    if (nsteps < 0) return bfail(b, SDT_ERR_ARG, "nsteps must be >= 0");
    const int n = (int)b->shots.size();
    std::vector<int> adv((size_t)n);
    int todo = 0;
    for (int q = 0; q < n; ++q) {
        sdt_handle *h = b->shots[(size_t)q];
        if (h->failed) return bfail(b, SDT_ERR_STATE, "a batch handle has already blown up");
        const int left = h->d.time_steps - h->steps_done;
        adv[(size_t)q] = nsteps < left ? nsteps : left;
        todo = adv[(size_t)q] > todo ? adv[(size_t)q] : todo;
    }
    hipError_t e = hipSetDevice(b->shots[0]->p.device);
    if (e != hipSuccess) return bfail(b, SDT_ERR_HIP, hipGetErrorString(e));
    Batch &B = b->b;
    // checkpointing shots (adjoint) advance in lockstep; pieces end at
    // multiples of the interval and at each shot's last step, where the
    // states are saved
    const int every = b->shots[0]->d.adj.every;
    for (int q = 0; q < n; ++q) {
        const sdt_handle *h = b->shots[(size_t)q];
        if (h->d.adj.every != every || (every > 0 && h->steps_done != b->shots[0]->steps_done))
            return bfail(b, SDT_ERR_STATE, "checkpointing batch shots need one interval and one step count");
    }
    int rc = SDT_OK;
    if (every <= 0) {
        rc = batch_advance(b, todo);
    } else {
        const int pos = b->shots[0]->steps_done;
        for (int done = 0; done < todo && rc == SDT_OK;) {
            const int p = pos + done;
            int piece = todo - done;
            piece = piece < every - p % every ? piece : every - p % every;
            for (int q = 0; q < n; ++q) {
                Device &d = b->shots[(size_t)q]->d;
                if (p >= d.time_steps) continue;
                if (p % every == 0) checkpoint_copy(d, p / every, true, B.stream);
                piece = piece < d.time_steps - p ? piece : d.time_steps - p;
            }
            rc = batch_advance(b, piece);
            done += piece;
            for (int q = 0; q < n; ++q) {
                Device &d = b->shots[(size_t)q]->d;
                if (pos + done == d.time_steps) checkpoint_copy(d, d.adj.nckpt - 1, true, B.stream);
            }
        }
    }
    if (rc) return rc;
    if ((e = hipGetLastError()) != hipSuccess) return bfail(b, SDT_ERR_HIP, hipGetErrorString(e));
    if ((e = hipStreamSynchronize(B.stream)) != hipSuccess) return bfail(b, SDT_ERR_HIP, hipGetErrorString(e));
    // every shot is checked and advanced; the first blow-up is reported
    int status = SDT_OK;
    for (int q = 0; q < n; ++q) {
        sdt_handle *h = b->shots[(size_t)q];
        if (adv[(size_t)q] == 0) continue;
        const int first = h->steps_done + 1;
        int32_t fs = 0;
        int rc = check_norms(h, first, adv[(size_t)q], &fs);
        if (rc == SDT_ERR_UNSTABLE) {
            if (status == SDT_OK) {
                if (failed_shot) *failed_shot = q;
                if (failed_step) *failed_step = fs;
                status = bfail(b, rc, "shot " + std::to_string(q) + ": " + h->err);
            }
            continue;
        }
        if (rc) return bfail(b, rc, h->err);
        h->steps_done += adv[(size_t)q];
    }
    if (status != SDT_OK) g_err = b->err;   // later shots' checks overwrote it
    return status;
}

int32_t sdt_batch_info(const sdt_batch *b, int32_t key, int64_t *value)
{
    if (!b || !value) return fail(nullptr, SDT_ERR_ARG, "null argument");
    switch (key) {
    case SDT_BATCH_INFO_TILE_ROWS:
        *value = b->b.choice < 0 ? 0 : (int64_t)kTile2dWaves * kBatchShapes[b->b.choice].rows;
        return SDT_OK;
    case SDT_BATCH_INFO_XCD_ORDER:
        *value = b->b.choice < 0 ? -1 : (int64_t)kBatchShapes[b->b.choice].xcd_order;
        return SDT_OK;
    case SDT_BATCH_INFO_TILES: *value = b->b.ntiles[b->b.choice < 0 ? 0 : b->b.choice]; return SDT_OK;
    }
    return fail(nullptr, SDT_ERR_ARG, "unknown batch info key");
}

int32_t sdt_batch_destroy(sdt_batch *b)
{
    if (!b) return SDT_OK;
    (void)hipSetDevice(b->shots[0]->p.device);
    (void)hipStreamSynchronize(b->b.stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (auto ev : b->ev)
        if (ev) (void)hipEventDestroy(ev);
    batch_free(b->b);
    for (sdt_handle *h : b->shots) h->owner = nullptr;
    delete b;
    return SDT_OK;
}

int32_t sdt_batch_adjoint(sdt_batch *b, const double *const *trace_grad, double *const *coef_grad,
                          double *const *src_grad)
{
    if (!b || !trace_grad || !coef_grad) return fail(nullptr, SDT_ERR_ARG, "null argument");
    const int n = (int)b->shots.size();
    sdt_handle *h0 = b->shots[0];
    const int every = h0->d.adj.every, nt = h0->d.time_steps;
    for (int q = 0; q < n; ++q) {
        const sdt_handle *h = b->shots[(size_t)q];
        if (h->d.adj.every <= 0 || h->d.adj.every != every || h->d.time_steps != nt || h->steps_done != nt)
            return bfail(b, SDT_ERR_STATE,
                         "every shot must have run all of one time_steps with one SDT_OPT_CHECKPOINT interval");
        if (h->failed) return bfail(b, SDT_ERR_STATE, "a batch handle has blown up");
        if (!coef_grad[q] || (h->d.nrec > 0 && !trace_grad[q])) return bfail(b, SDT_ERR_ARG, "null per-shot array");
    }
    hipError_t e = hipSetDevice(h0->p.device);
    if (e != hipSuccess) return bfail(b, SDT_ERR_HIP, hipGetErrorString(e));
    Batch &B = b->b;
    std::vector<Device *> devs((size_t)n);
    for (int q = 0; q < n; ++q) {
        sdt_handle *h = b->shots[(size_t)q];
        devs[(size_t)q] = &h->d;
        if (adjoint_alloc(h) != SDT_OK) return bfail(b, SDT_ERR_HIP, "adjoint allocation failed (out of device memory?)");
        int rc = adjoint_reset(h, trace_grad[q]);
        if (rc) return bfail(b, rc, h->err);
    }
    if (batch_adj_args(B, devs.data(), n) != SDT_OK) return bfail(b, SDT_ERR_HIP, "adjoint argument upload failed");
    const int w = B.choice < 0 ? 1 : B.choice;
    const int nseg = h0->d.adj.nckpt - 1;
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int first = seg * every;
        const int len = nt - first < every ? nt - first : every;
        for (int q = 0; q < n; ++q) {
            Device &d = *devs[(size_t)q];
            d.adj.store_first = first;
            checkpoint_copy(d, seg, false, B.stream);
            set_step_counter(B.stream, d.d_it, first + 1);
        }
        if (batch_refresh_args(B, devs.data(), n) != SDT_OK) return bfail(b, SDT_ERR_HIP, "argument upload failed");
        B.store = true;
        for (int s = 0; s < len; ++s) batch_step(B, w);
        B.store = false;
        for (int t = first + len - 1; t >= first; --t) batch_adj_step(B, w, t, first);
        batch_adj_correlate(B, w, len);
        if ((e = hipGetLastError()) != hipSuccess) return bfail(b, SDT_ERR_HIP, hipGetErrorString(e));
    }
    if (B.physics == SDT_ELASTIC2) el2_batch_adj_finish(B, w);
    for (int q = 0; q < n; ++q) {   // back to the forward run's final states
        Device &d = *devs[(size_t)q];
        checkpoint_copy(d, d.adj.nckpt - 1, false, B.stream);
        set_step_counter(B.stream, d.d_it, nt + 1);
        d.adj.store_first = 0;
    }
    if ((e = hipStreamSynchronize(B.stream)) != hipSuccess) return bfail(b, SDT_ERR_HIP, hipGetErrorString(e));
    if (batch_refresh_args(B, devs.data(), n) != SDT_OK) return bfail(b, SDT_ERR_HIP, "argument upload failed");
    for (int q = 0; q < n; ++q) {
        int rc = adjoint_download(b->shots[(size_t)q], coef_grad[q], src_grad ? src_grad[q] : nullptr);
        if (rc) return bfail(b, rc, b->shots[(size_t)q]->err);
    }
    return SDT_OK;
}

int32_t sdt_synchronize(sdt_handle *h)
{
    if (!h) return fail(nullptr, SDT_ERR_ARG, "null handle");
    HIPCHK(h, hipSetDevice(h->p.device));
    HIPCHK(h, hipStreamSynchronize(h->d.stream));
    return SDT_OK;
}

}  // extern "C"
