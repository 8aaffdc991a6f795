// sdt_internal.h -- host-side state shared between the C-ABI (sdt_api.cpp)
// and the per-solver launchers (*.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>

#include "../../include/seidart_hip.h"
#include "host_layout.h"

namespace sdt {

constexpr int kMaxMem = 32;     // memory-variable arrays per solver (3D elastic: 29)
constexpr int kAltMem = 16;     // 3D EM fused step: second buffers of mem[0..5] start here
constexpr int kMaxMat = 32;     // device material arrays per solver (3D elastic: 28)
constexpr double kEps0 = 8.85418782e-12;                  // constants.f08:10
constexpr double kMu0 = 4.0 * 3.141592653589793238462643 * 1.0e-7;  // constants.f08:9
constexpr double kClight = 2.9979458e8;                    // constants.f08:8
constexpr double kStability = 1.0e25;                    // constants.f08:12

// One point-source injection: field[cell] += (series[it] * mul) / div.
struct SrcTerm {
    int64_t cell;     // linear index, -1 = inactive
    int field;        // SDT_F_*
    int component;    // SDT_SRC_*
    double mul, div;  // reference evaluation order: (src * mul) / div
    int on_face;      // velocity/E injection on a Dirichlet face: dropped
};

// Cells whose norm contribution the tail kernel recomputes after injection.
constexpr int kMaxNormFix = 9;

// FP64 window around the point source of a 2D EM shot run in FP32: its
// cells' H and E updates are done again in FP64 every step (FP64 values,
// FP64 coefficients), the source is added there in FP64, the FP32 fields
// receive the rounded values, and receivers inside record the FP64 values.
// Near the source the fields are large and FP32 rounding of them swamps the
// small differences between two models that an FWI residual is made of.
// The window lies in the CPML-free box, off the Dirichlet faces.
constexpr int kSrcWinHalf = 7;                     // cells on each side of the source
constexpr int kSrcWinMax = 2 * kSrcWinHalf + 1;    // window side at most (one thread per cell)
struct SrcWindow {
    double *v;           // null: no window; else [3][(ni + 1) * (nj + 1)] Ex, Ez, Hy of the
                         // window, its ring column i0 + ni and ring row j0 + nj (E only)
    const double *co;    // [7][ni * nj] update coefficients aEx bEx bEz dEx eEx dEz eEz
    int i0, j0, ni, nj;  // first cell (1-based) and size
    int pitch;
    int64_t base;        // padded cell index of (i0, j0)
    double rdx, rdz, dt, dbHy;
};

struct TailArgs {
    int *d_it;                       // device step counter (1-based step being run)
    const double *series;            // [SDT_NUM_SRC][time_steps]
    int time_steps;
    SrcTerm terms[SDT_NUM_SRC];
    int nterms;
    int64_t vfix[kMaxNormFix];       // cells excluded from the fused velocity norm
    int nvfix;
    int64_t sfix[kMaxNormFix];       // cells excluded from the fused stress norm
    int nsfix;
    int nvcomp;                      // components in the velocity norm (2 or 3)
    int vcomp[3];                    // their field ids
    int stress_norm;                 // 1 for 3D elastic
    unsigned long long *vnorm;       // [time_steps] bit patterns
    unsigned long long *snorm;       // [time_steps]
    int nrec;
    const int64_t *rec_cell;         // [nrec]
    double *rec_out;                 // [time_steps][nrec][3]
    int rec_field[3];                // field ids recorded (-1 -> 0.0)
    int cplx;                        // 1: complex fields (interleaved re, im)
    SrcWindow win;                   // FP64 source window (win.v null: none)
};

// Adjoint-state gradients (2D EM, 2D elastic).  The forward run saves its
// state every `every` steps; the adjoint run walks the segments backwards:
// it restores a segment's first state, recomputes the segment while the
// forward kernels store the values the gradient needs (seg_fwd), steps the
// adjoint fields back through the segment (storing the adjoint values the
// gradient needs in seg_adj), then correlates the two per cell into the FP64
// gradient accumulators.
constexpr int kMaxAdjArrays = 32;
constexpr int kAdjSrcSlots = 8;        // adjoint values recorded per step at the source cells
struct Adjoint {
    int every = 0;                     // SDT_OPT_CHECKPOINT (0 = no checkpoints)
    int nckpt = 0;                     // slot k = state after min(k*every, time_steps) steps
    int nstate = 0;                    // arrays per checkpoint
    void *state[kMaxAdjArrays] = {};   // the device arrays a checkpoint holds
    char *ckpt = nullptr;              // [nckpt][nstate][ncell_dev] elements of T
    int nfwd = 0, nadj = 0;            // values per cell stored per step
    void *seg_fwd = nullptr;           // [every][nfwd][ncell_dev]
    void *seg_adj = nullptr;           // [every][nadj][ncell_dev]
    void *lam[kMaxAdjArrays] = {};     // adjoint fields, memory variables, transfer arrays
    int nlam = 0;
    int ncoef = 0;                     // update coefficients per cell
    int ngrad = 0;                     // FP64 accumulator arrays (>= ncoef; the first ncoef are the gradient)
    double *grad = nullptr;            // [ngrad][ncell_dev]
    double *src_lam = nullptr;         // [time_steps][kAdjSrcSlots] adjoint values at the source cells
    double *trace_grad = nullptr;      // [time_steps][nrec][3]
    double *win_ckpt = nullptr;        // [nckpt] copies of the source window's values
    bool store = false;                // forward kernels write seg_fwd
    int store_first = 0;               // 0-based first step of the recomputed segment
};

// Plane-wave source: the cells of the active interior faces (face[a] = cell
// index of the face normal to axis a, 0 = no face; lo/hi = cell ranges of
// the other axes).  The fused norm reductions skip these cells; the
// plane-wave pass folds their post-injection values in.
struct PwFaces {
    int face[3];
    int lo[3], hi[3];
};
__host__ __device__ inline bool pw_face(const PwFaces &f, int i, int j, int k)
{
    return (i == f.face[0] && j >= f.lo[1] && j <= f.hi[1] && k >= f.lo[2] && k <= f.hi[2]) ||
           (j == f.face[1] && i >= f.lo[0] && i <= f.hi[0] && k >= f.lo[2] && k <= f.hi[2]) ||
           (k == f.face[2] && i >= f.lo[0] && i <= f.hi[0] && j >= f.lo[1] && j <= f.hi[1]);
}

struct PlaneWaveSpec {   // sdt_set_plane_wave arguments
    int kind = 0;        // 0 none, SDT_PW_S, SDT_PW_P
    double angle_xz = 0, angle_xy = 0, angle_yz = 0, f0 = 0, amplitude = 0;
    int wavelet = 0;
};

struct PlaneWave {
    int on = 0;
    int face[3] = {}, lo[3] = {}, hi[3] = {};
    double p[3] = {}, e[3] = {};   // propagation direction, polarisation
    double r0[3] = {};             // reference point (m)
    double loc[3] = {};            // face positions (m)
    double h[3] = {};              // cell sizes
    double v = 0, f0 = 0, amp = 0, dt = 0;   // background speed, wavelet, step
    int wavelet = 0;
    int em = 0;                    // 1: electromagnetic (E along e, H along hh)
    double eta = 0;                // background impedance (ohm)
    double hh[3] = {};             // unit p x e
    PwFaces faces() const
    {
        PwFaces f;
        for (int a = 0; a < 3; ++a) {
            f.face[a] = on ? face[a] : 0;
            f.lo[a] = lo[a];
            f.hi[a] = hi[a];
        }
        return f;
    }
};

// trace-gradient injection of one adjoint step
struct AdjTailArgs {
    void *lam[3];             // adjoint array of each recorded component (or null)
    const int64_t *rec_cell;
    const double *tg;         // [time_steps][nrec][3]
    int nrec;
};

struct Device {
    int physics = 0, precision = 0;
    int nx = 0, ny = 0, nz = 0;
    double dx = 0, dy = 0, dz = 0, dt = 0;
    int time_steps = 0;
    int isrc = 0, jsrc = 0, ksrc = 0;   // padded 1-based
    int cpml_width = 0;                // absorbing-layer width (cells)
    int density_code = 3, stencil = 0;
    double ky = 0.0;                   // 2.5D elastic y wavenumber
    int ncomp = 1;                     // values per field element (2 = complex)
    int64_t ncell = 0;
    size_t esize = 8;                  // sizeof(T)
    // Device layout of grid arrays: row (j, k) holds x cells 1..nx at
    // elements xoff .. xoff+nx-1 of a row of `pitch` elements; pitch is a
    // multiple of a 128-byte cache line and xoff puts the first interior
    // column on a line boundary.  Cell (i, j, k) -> xoff + (i-1) +
    // pitch * ((j-1) + ny * (k-1)).
    int pitch = 0;
    int64_t xoff = 0;
    int64_t ncell_dev = 0;

    void *field[SDT_NUM_FIELDS] = {};
    // 3D EM with the fused step: the other buffer of each of E and H (the
    // step reads field[] and writes alt[], then the two are swapped), and of
    // the H-update memory variables mem[0..5] in mem[kAltMem + 0..5]
    void *alt[SDT_NUM_FIELDS] = {};
    int parity = 0;                    // buffer swaps so far, mod 2
    int fused_opt = -1;                // SDT_OPT_FUSED (-1 auto)
    bool fused = false;                // fused 3D EM step in use
    void *mem[kMaxMem] = {};
    int nmem = 0;
    void *mat[kMaxMat] = {};
    int nmat = 0;
    void *cpml[6] = {};                // a, b, 1/kappa (full), a, b, 1/kappa (half)
    // or, when the grid holds at most 256 distinct such tuples: a one-byte id
    // per padded cell and a table of rows of 8 (6 used) (upload_cpml_dict)
    uint8_t *cpml_cid = nullptr;
    void *cpml_ctab = nullptr;
    int cpml_ncoef = 0;
    uint8_t *tile_flags = nullptr;     // 1 = tile holds a CPML cell inside the interior box
    int64_t nflagged = 0;              // tiles flagged inside the box
    void *tiles = nullptr;             // 3D elastic CPML-shell tile list (device)
    int ntiles = 0;
    int64_t nslots = 0;                // compact CPML storage length (cells x planes)
    void *shell = nullptr;             // 3D elastic blocked shell storage (a grid allocation)
    int shell_snq = 0;                 // streams per shell tile-plane block
    uint8_t *shell_cid = nullptr;      // CPML coefficient id per shell slot (or null)
    void *shell_ctab = nullptr;        // CPML coefficient rows (a grid allocation)
    int shell_ncoef = 0;               // distinct coefficient tuples (0 = in the block)
    int box[6] = {1, 0, 1, 0, 1, 0};   // interior box ilo,ihi,jlo,jhi,klo,khi (1-based, inclusive):
                                       // every cell inside has identity CPML coefficients
    int tiles_x = 0, tiles_y = 0, tiles_z = 0;
    int tile_x = 64, tile_y = 4, tile_z = 1;

    double *series = nullptr;          // [SDT_NUM_SRC][time_steps]
    unsigned long long *norms = nullptr;  // [2][time_steps]
    int *d_it = nullptr;
    int64_t *rec_cell = nullptr;
    double *rec_out = nullptr;
    int nrec = 0;

    TailArgs tail{};
    int zchunk = 0;                    // SDT_OPT_ZCHUNK
    int mat_table_opt = 1;             // SDT_OPT_MATERIAL_TABLE
    bool mat_table = false;            // per-cell material ids + table in use
    int iso_opt = 1;                   // SDT_OPT_ISOTROPIC
    bool mat_iso = false;              // 3D elastic: only c11, c12, c44 held (isotropic form)
    bool ig_same = false;              // 3D elastic per-cell arrays: the six attenuation divisors are equal
    int nmaterials = 0;                // distinct materials found at setup
    uint8_t *mid = nullptr;            // material id per (padded) cell
    void *tab = nullptr;               // material rows
    int tile_rows = 0;                 // SDT_OPT_TILE_ROWS
    bool graphs = true;                // SDT_OPT_GRAPH
    bool stagger = true;               // SDT_OPT_STAGGER
    int nalloc = 0;                    // allocations made (stagger index)
    void *bases[256] = {};             // hipMalloc'd base of every array
    int nbases = 0;
    size_t grid_bytes = 0;             // bytes allocated through alloc_grid
    hipStream_t stream = nullptr;
    // second stream for the 3D elastic CPML-shell kernels, which run
    // concurrently with the interior kernels of the same pass (fork/join by
    // events); both streams have the default priority
    hipStream_t side = nullptr;
    hipEvent_t fork[2] = {}, join[2] = {};
    bool concurrent = false;           // SDT_OPT_CONCURRENT
    int xcd_remap = 0;                 // SDT_OPT_XCD_REMAP
    hipEvent_t ev[4] = {};
    bool timing = false;
    double ms_stress = 0, ms_vel = 0;
    int64_t timed_launches = 0;
    void *snap = nullptr;                // snapshot staging (ncell doubles), allocated on first use

    const double *coef_in = nullptr;   // update coefficients given directly (during setup only)
    bool raw_source = false;           // SDT_OPT_RAW_SOURCE: series injected as given
    bool src_window_opt = true;        // SDT_OPT_SOURCE_WINDOW (2D EM in FP32)
    Adjoint adj;
    PlaneWave pw;                      // plane-wave source (elastic solvers)
};

// Per-solver hooks, defined in the solver's .hip file.
//   setup: convert the staged FP64 inputs (host_in) into the solver's device
//          layout; called once before the first step.
//   step:  enqueue one time step (stress/H kernel, velocity/E kernel, tail).
typedef int (*setup_fn)(Device &d, const double *const *host_in);
typedef void (*step_fn)(Device &d, bool timed);

int el2_setup(Device &d, const double *const *host_in);
void el2_step(Device &d, bool timed);
int el25_setup(Device &d, const double *const *host_in);
void el25_step(Device &d, bool timed);
int el3_setup(Device &d, const double *const *host_in);
void el3_step(Device &d, bool timed);
int em25_setup(Device &d, const double *const *host_in);
void em25_step(Device &d, bool timed);
int em2_setup(Device &d, const double *const *host_in);
void em2_step(Device &d, bool timed);
int em3_setup(Device &d, const double *const *host_in);
void em3_step(Device &d, bool timed);

// adjoint hooks of the 2D solvers: allocate the adjoint arrays, reverse step
// t (0-based; its trace row is injected first), correlate the stored
// segment [first, first + n) into the gradient
int em2_adj_setup(Device &d);
void em2_adj_step(Device &d, int t);
void em2_adj_correlate(Device &d, int first, int n);
int el2_adj_setup(Device &d);
void el2_adj_step(Device &d, int t);
void el2_adj_correlate(Device &d, int first, int n);
void el2_adj_finish(Device &d);   // density gradient from its accumulated sums

// Many 2D shots advanced together: one launch per pass covers the tiles of
// every shot (tile list entries name the shot), one tail launch covers every
// shot's sources and receivers.  A shot whose step counter passed its
// time_steps is skipped, so shots of different lengths share the launches.
// 2D tiles: 64 line-aligned columns x kTile2dWaves wavefronts, each marching
// R consecutive rows; single-shot launches use R = kSingleRows2d.  A batch
// times each of kBatchShapes (rows per wavefront; tile-list order) on its
// first steps and keeps the fastest.  In XCD order the tile list is permuted
// so that runs of 8 x-consecutive tiles land on one XCD (workgroups are
// dealt to the 8 XCDs round-robin by index), letting the column halos of a
// tile hit in that XCD's L2.
constexpr int kTile2dWaves = 4;
constexpr int kSingleRows2d = 4;
struct BatchShape {
    int rows;
    bool xcd_order;
};
constexpr int kNumBatchShapes = 4;
constexpr BatchShape kBatchShapes[kNumBatchShapes] = {{4, false}, {8, false}, {4, true}, {8, true}};
inline int tiles_x_2d(const Device &d) { return (int)((d.xoff + d.nx + 63) / 64); }
inline int tiles_z_2d(const Device &d, int rows) { return (d.nz + kTile2dWaves * rows - 1) / (kTile2dWaves * rows); }
// A 2D elastic single-shot grid too small to give every CU a few
// workgroups at kSingleRows2d rows per wavefront (the 220 x 220 case gives
// 56) runs one row per wavefront: four times the workgroups, a quarter of
// the serial row latency per wavefront; same per-cell arithmetic either way.
// (Measured on 64 shots of 300-800 x 150-400: 2D elastic 7.4 -> 8.8
// GCell/s, 2D EM 10.8 -> 10.1, so the 2D EM updates keep kSingleRows2d.)
constexpr int kSmallGrid2dTiles = 1024;
inline bool small_grid_2d(const Device &d) { return tiles_x_2d(d) * tiles_z_2d(d, kSingleRows2d) < kSmallGrid2dTiles; }
struct Batch {
    int physics = 0, precision = 0, nshots = 0;
    void *args = nullptr;          // per-shot kernel arguments (device)
    int4 *tiles[kNumBatchShapes] = {};   // per shape: (shot or -1, tile x, band, 0)
    int ntiles[kNumBatchShapes] = {};
    void *fps = nullptr;           // per-shot field pointers (device)
    TailArgs *tails = nullptr;     // per-shot tail arguments (device)
    int choice = -1;               // autotuned tile shape index, -1 = not yet
    bool store = false;            // forward E / velocity kernels write the segment store
    bool any_table = false;        // some shot reads a material table (else the kernels without it)
    void *adj_args = nullptr;      // per-shot adjoint kernel arguments (device)
    AdjTailArgs *adj_tails = nullptr;
    hipStream_t stream = nullptr;
};
int batch_setup(Batch &b, Device *const *shots, int n);
void batch_free(Batch &b);
// one step of every unfinished shot with tile shape kBatchShapes[which]
void batch_step(Batch &b, int which);
int el2_batch_args(Batch &b, Device *const *shots, int n);
int em2_batch_args(Batch &b, Device *const *shots, int n);
void el2_batch_launch(Batch &b, int which);
void em2_batch_launch(Batch &b, int which);
void launch_tail_batch(Batch &b);
// batched adjoint: (re)build the per-shot forward arguments (after changing
// a shot's segment start) and adjoint arguments, reverse step t of every
// shot, correlate the stored segment of every shot
int batch_refresh_args(Batch &b, Device *const *shots, int n);
int batch_adj_args(Batch &b, Device *const *shots, int n);
void batch_adj_step(Batch &b, int which, int t, int first);
void batch_adj_correlate(Batch &b, int which, int n);
int em2_batch_adj_args(Batch &b, Device *const *shots, int n);
void em2_batch_adj_step(Batch &b, int which, int t, int first);
void em2_batch_adj_correlate(Batch &b, int which, int n);
int el2_batch_adj_args(Batch &b, Device *const *shots, int n);
void el2_batch_adj_step(Batch &b, int which, int t, int first);
void el2_batch_adj_correlate(Batch &b, int which, int n);
void el2_batch_adj_finish(Batch &b, int which);
void launch_adj_tail_batch(Batch &b, int t);

// common pieces (sdt_common.hip)
void launch_tail(Device &d);
// plane-wave source: geometry and background speed from the staged model
// (before its arrays are released), and the per-step injection
int plane_wave_setup(Device &d, const PlaneWaveSpec &spec, const double *const *host_in);
// every cell's stiffness (c11..c66 inputs, absent = 0) has the isotropic form
bool stiffness_isotropic(const double *const *host_in, int64_t n);
void launch_plane_wave(Device &d);
constexpr int kEl3TabRow = 32;         // 3D elastic material-table row length
constexpr int kEl25TabRow = 32;        // 2.5D elastic material-table row length
constexpr int kEl2TabRow = 32;         // 2D elastic material-table row length
constexpr int kEm2TabRow = 8;          // 2D EM material-table row length
constexpr int kEm25TabRow = 16;        // 2.5D EM material-table row length
void launch_adj_tail(Device &d, int t, void *const lam_field[3]);
void set_step_counter(hipStream_t s, int *d_it, int v);
int download_grid_f64(Device &d, const double *dev, double *host);
// snapshot pack: planes [k0, k0 + nk) of a device grid array -> Fortran order
// in dst (FP32 when single, else FP64); zero_faces reads the Dirichlet faces
// of a lazily zeroed H field as 0
void launch_pack(const Device &d, const void *src, void *dst, int single, int k0, int nk, bool zero_faces,
                 hipStream_t s);
// CPML classification.  A cell is an identity cell when a == 0, kappa == 1
// and b is finite in both the full and the half coefficient set (its memory
// variables then stay exactly zero and its derivatives pass unchanged).
// d.box is the box grown from the grid centre along the three centre lines
// while cells stay identity cells; cells outside it take the CPML body.
// tile_flags marks the (tile_x, tile_y, tile_z) tiles that nevertheless hold
// a non-identity cell inside the box, whose cells then take the CPML body
// too.  Kernels test: cpml(cell) = !inside(box) || tile_flags[tile].
int compute_tile_flags(Device &d, const double *const *host_in);
// full-grid CPML coefficient arrays: a, b, 1/kappa for the full and half
// sets (identity defaults where not supplied)
int upload_cpml_full(Device &d, const double *const *host_in);
// The same coefficients as a dictionary (cpml_cid, cpml_ctab) when there are
// at most 256 distinct tuples, otherwise as upload_cpml_full
int upload_cpml_dict(Device &d, const double *const *host_in);
// choose pitch / xoff (call after compute_tile_flags): column `first` starts
// a 128-byte line; 0 = the interior box's first column (d.box)
void set_layout(Device &d, int first = 0);
// source window values (bytes of win.v) and their refresh from an unpadded
// FP64 host field (Fortran order) after the field is set
size_t src_window_bytes(const SrcWindow &w);
int src_window_load(Device &d, int f, const double *host);
inline int64_t cell_index(const Device &d, int i, int j, int k)
{
    return d.xoff + (int64_t)(i - 1) + (int64_t)d.pitch * ((int64_t)(j - 1) + (int64_t)d.ny * (k - 1));
}
// grid arrays in the padded layout: upload an FP64 host array given in
// Fortran order (nx, ny, nz), or allocate a zeroed one; ncomp = 2 for complex
// elements (interleaved re, im on host and device)
int upload_grid(Device &d, const double *host, void **dev, int ncomp = 1);
int alloc_grid_zeroed(Device &d, void **dev, int ncomp = 1);
// copy a device grid array to / from an FP64 host array in Fortran order
int download_grid(Device &d, const void *dev, double *host, int ncomp = 1);
int store_grid(Device &d, void *dev, const double *host, int ncomp = 1);
// copy n FP64 host values into a freshly allocated device array of the
// handle's precision
int upload_converted(Device &d, const double *host, void **dev, int64_t n);
int alloc_zeroed(Device &d, void **dev, int64_t n);
// hipMalloc of a grid array whose start is offset by a different multiple of
// 4352 bytes per allocation (when d.stagger), so that the dozens of equally
// sized arrays streamed together do not begin on the same HBM channel.
int alloc_grid(Device &d, void **dev, size_t bytes);
void free_grids(Device &d);
// Material dictionary: fill(c, row) writes the row_len material values of
// host cell c (unused entries stay 0).  When the grid holds at most 256
// distinct rows, d.mid (one byte per padded cell), d.tab (rows of row_len in
// the handle's precision), d.nmaterials and d.mat_table are set; otherwise
// nothing is allocated and d.mat_table stays false.
int build_material_table(Device &d, int row_len, const std::function<void(int64_t, double *)> &fill);

}  // namespace sdt
