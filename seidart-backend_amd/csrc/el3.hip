// el3.hip -- 3D anisotropic elastic CPML-FDTD (seismic3,
// cpmlfdtd.f08:1341-2313) on gfx950.
//
// Setup, device layout and the step sequence.  The reference's four stress
// nests (L1 sigma_xx/yy/zz, L2 sigma_xy, L3 sigma_xz, L4 sigma_yz;
// cpmlfdtd.f08:1761-1992) run fused in one stress pass and its three velocity
// nests (2014-2127) in one velocity pass that also applies the Dirichlet
// faces (2226-2248) and the velocity norm (2255); each pass is an interior
// kernel plus a CPML-shell kernel (el3_fast.hip).
//
// Memory variables are deduplicated: a recursion psi <- b psi + a d depends
// only on its own cell's history of d, so the reference arrays that share
// (velocity component, direction, one-sided stencil, coefficient set) hold
// identical values wherever they are updated, and one array updated on the
// union of their ranges is exact.  Stress side, per component: x {D+ half,
// D- full}, y {D- full, D+ half}, z {D- full, D- half, D+ half} = 21 arrays;
// vx's (z, D-, half) recursion (memory_dvx_dz2, 1855) is written but never
// read by the reference (1858-1865) and is not kept.  Velocity side: 9.
#include <math.h>
#include <string.h>

#include <unordered_map>
#include <vector>

#include "el3_common.h"

namespace sdt {

template <typename T>
static El3Args<T> el3_args(Device &d)
{
    El3Args<T> A;
    for (int q = 0; q < 3; ++q) A.v[q] = (T *)d.field[SDT_F_VX + q];
    for (int q = 0; q < 6; ++q) A.s[q] = (T *)d.field[SDT_F_SXX + q];
    for (int q = 0; q < 21; ++q) A.C[q] = (const T *)d.mat[q];
    A.rho = (const T *)d.mat[21];
    for (int q = 0; q < 6; ++q) A.ig[q] = (const T *)d.mat[22 + q];
    A.mid = d.mid;
    A.tab = (const T *)d.tab;
    A.shell = (T *)d.shell;
    A.remap = d.xcd_remap;
    A.ig_same = d.ig_same ? 1 : 0;
    A.edge = d.concurrent ? 0 : 1;
    A.snq = d.shell_snq;
    A.nslots = d.nslots;
    A.cid = d.shell_cid;
    A.ctab = (const T *)d.shell_ctab;
    A.flags = d.tile_flags;
    for (int q = 0; q < 6; ++q) A.box[q] = d.box[q];
    A.nx = d.nx; A.ny = d.ny; A.nz = d.nz;
    A.pitch = d.pitch; A.xoff = d.xoff;
    A.tiles_x = d.tiles_x; A.tiles_y = d.tiles_y;
    A.rd24[0] = (T)(1.0 / (24.0 * d.dx));
    A.rd24[1] = (T)(1.0 / (24.0 * d.dy));
    A.rd24[2] = (T)(1.0 / (24.0 * d.dz));
    A.dt = (T)d.dt;
    A.density_code = d.density_code;
    A.corrected = d.stencil == SDT_STENCIL_CORRECTED;
    A.d_it = d.d_it;
    A.vnorm = d.norms;
    A.snorm = d.norms + d.time_steps;
    for (int q = 0; q < 3; ++q) A.vfix[q] = d.tail.vfix[q];
    A.pwf = d.pw.faces();
    for (int q = 0; q < 4; ++q) A.sfix[q] = d.tail.sfix[q];
    return A;
}

template <typename T>
static void el3_step_t(Device &d, bool timed)
{
    const El3Args<T> A = el3_args<T>(d);
    if (timed) (void)hipEventRecord(d.ev[0], d.stream);
    el3_fast_stress<T>(d, A);
    if (timed) (void)hipEventRecord(d.ev[1], d.stream);
    el3_fast_velocity<T>(d, A);
    if (timed) (void)hipEventRecord(d.ev[2], d.stream);
    launch_plane_wave(d);
    launch_tail(d);
}

void el3_step(Device &d, bool timed)
{
    if (d.precision == SDT_FP64) el3_step_t<double>(d, timed);
    else el3_step_t<float>(d, timed);
}

int el3_build_shell(Device &d, const double *const *in, int kc);

// Material dictionary.  A cell's material is the tuple (c11..c66,
// 1/(1+gamma*dt) for x y z xy xz yz, rho) with the same FP64 values the
// per-cell arrays would hold.  When the model has at most 256 distinct tuples
// (layered and blocky models, as SeidarT builds them from a material image),
// the kernels read a one-byte id per cell and the tuple from a small table
// that stays in cache, instead of 28 per-cell arrays; the values are the same.
// Otherwise d.mat_table stays false and the per-cell arrays are used.
static int el3_material_table(Device &d, const double *const *in)
{
    const int gids[] = {SDT_GAMMA_X, SDT_GAMMA_Y, SDT_GAMMA_Z, SDT_GAMMA_XY, SDT_GAMMA_XZ, SDT_GAMMA_YZ};
    return build_material_table(d, kTabRow, [&](int64_t c, double *r) {
        for (int q = 0; q < 21; ++q) r[q] = in[SDT_C11 + q] ? in[SDT_C11 + q][c] : 0.0;
        for (int q = 0; q < 6; ++q) r[kTabIg + q] = 1.0 / (1.0 + (in[gids[q]] ? in[gids[q]][c] : 0.0) * d.dt);
        r[kTabRho] = in[SDT_RHO] ? in[SDT_RHO][c] : 0.0;
    });
}

int el3_setup(Device &d, const double *const *in)
{
    d.tile_x = 64;
    d.tile_y = 4;
    d.tile_z = 1;
    int rc = compute_tile_flags(d, in);
    if (rc) return rc;
    host::el3_interior_box(d.box, d.nflagged, d.nx, d.ny, d.nz);
    set_layout(d);

    const int64_t n = d.ncell;
    std::vector<double> zero;
    d.nmat = 0;
    rc = d.mat_table_opt ? el3_material_table(d, in) : SDT_OK;
    if (rc) return rc;
    d.mat_iso = !d.mat_table && d.iso_opt && stiffness_isotropic(in, n);
    if (d.mat_iso) {
        // c11, c12, c44 in their Voigt slots 0, 1, 15; the other 18 slots stay null
        for (int id : {SDT_C11, SDT_C12, SDT_C44, SDT_RHO}) {
            if (!in[id] && zero.empty()) zero.assign((size_t)n, 0.0);
            rc = upload_grid(d, in[id] ? in[id] : zero.data(), &d.mat[id - SDT_C11]);
            if (rc) return rc;
        }
        d.nmat = SDT_RHO - SDT_C11 + 1;
    } else if (!d.mat_table) {
        for (int id = SDT_C11; id <= SDT_RHO; ++id) {
            if (!in[id] && zero.empty()) zero.assign((size_t)n, 0.0);
            rc = upload_grid(d, in[id] ? in[id] : zero.data(), &d.mat[d.nmat++]);
            if (rc) return rc;
        }
    }
    d.ig_same = false;
    if (!d.mat_table) {
        // attenuation divisors 1/(1 + gamma dt); a direction whose array
        // equals an earlier one's shares its device array, and when all six
        // are equal the stress kernels load one value per cell
        std::vector<double> inv[6];
        const int gids[] = {SDT_GAMMA_X, SDT_GAMMA_Y, SDT_GAMMA_Z, SDT_GAMMA_XY, SDT_GAMMA_XZ, SDT_GAMMA_YZ};
        int same = 0;
        for (int q = 0; q < 6; ++q) {
            const int id = gids[q];
            inv[q].resize((size_t)n);
            for (int64_t c = 0; c < n; ++c) inv[q][(size_t)c] = 1.0 / (1.0 + (in[id] ? in[id][c] : 0.0) * d.dt);
            int p = 0;   // an earlier array with the same values (aliased ones are left empty)
            while (p < q && (inv[p].empty() || memcmp(inv[p].data(), inv[q].data(), sizeof(double) * (size_t)n) != 0))
                ++p;
            if (p < q) {
                d.mat[d.nmat++] = d.mat[22 + p];
                std::vector<double>().swap(inv[q]);
                if (p == 0) ++same;
                continue;
            }
            rc = upload_grid(d, inv[q].data(), &d.mat[d.nmat++]);
            if (rc) return rc;
        }
        d.ig_same = same == 5;
    }
    for (int f = SDT_F_VX; f <= SDT_F_SYZ; ++f) {
        if (f <= SDT_F_VZ && in[SDT_IC_X + f]) rc = upload_grid(d, in[SDT_IC_X + f], &d.field[f]);
        else rc = alloc_grid_zeroed(d, &d.field[f]);
        if (rc) return rc;
    }

    // point sources, cpmlfdtd.f08:2208-2219
    const int is = d.isrc, js = d.jsrc, ks = d.ksrc, nx = d.nx, ny = d.ny, nz = d.nz;
    auto cell = [&](int i, int j, int k) { return cell_index(d, i, j, k); };
    auto face = [&](int i, int j, int k) {
        return (int)(i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz);
    };
    const double *rho = in[SDT_RHO];
    auto R = [&](int i, int j, int k) {
        return rho ? rho[(int64_t)(i - 1) + (int64_t)nx * ((int64_t)(j - 1) + (int64_t)ny * (k - 1))] : 0.0;
    };
    TailArgs &t = d.tail;
    t.nterms = 0;
    t.terms[t.nterms++] = {cell(is, js, ks), SDT_F_SXX, SDT_SRC_XX, 1.0, R(is, js, ks), 0};
    t.terms[t.nterms++] = {cell(is + 1, js + 1, ks), SDT_F_SXY, SDT_SRC_XY, 1.0, R(is + 1, js + 1, ks), 0};
    t.terms[t.nterms++] = {cell(is + 1, js, ks + 1), SDT_F_SXZ, SDT_SRC_XZ, 1.0, R(is + 1, js, ks), 0};
    t.terms[t.nterms++] = {cell(is, js, ks), SDT_F_SYY, SDT_SRC_YY, 1.0, R(is, js, ks), 0};
    t.terms[t.nterms++] = {cell(is, js + 1, ks + 1), SDT_F_SYZ, SDT_SRC_YZ, 1.0, R(is, js, ks + 1), 0};
    t.terms[t.nterms++] = {cell(is, js, ks), SDT_F_SZZ, SDT_SRC_ZZ, 1.0, R(is, js, ks), 0};
    t.terms[t.nterms++] = {cell(is + 1, js, ks), SDT_F_VX, SDT_SRC_X, d.dt, R(is + 1, js, ks), face(is + 1, js, ks)};
    t.terms[t.nterms++] = {cell(is, js + 1, ks), SDT_F_VY, SDT_SRC_Y, d.dt, R(is, js, ks), face(is, js + 1, ks)};
    t.terms[t.nterms++] = {cell(is, js, ks + 1), SDT_F_VZ, SDT_SRC_Z, d.dt, R(is, js, ks + 1), face(is, js, ks + 1)};
    t.nvfix = 3;
    t.vfix[0] = cell(is + 1, js, ks);
    t.vfix[1] = cell(is, js + 1, ks);
    t.vfix[2] = cell(is, js, ks + 1);
    t.nvcomp = 3;
    t.vcomp[0] = SDT_F_VX; t.vcomp[1] = SDT_F_VY; t.vcomp[2] = SDT_F_VZ;
    t.stress_norm = 1;
    t.nsfix = 4;
    t.sfix[0] = cell(is, js, ks);
    t.sfix[1] = cell(is + 1, js + 1, ks);
    t.sfix[2] = cell(is + 1, js, ks + 1);
    t.sfix[3] = cell(is, js + 1, ks + 1);
    t.rec_field[0] = SDT_F_VX; t.rec_field[1] = SDT_F_VY; t.rec_field[2] = SDT_F_VZ;

    // planes per shell tile: short chunks keep the planes that the resident
    // shell workgroups touch at one time close together (280^3 FP64: 1.891 ms
    // per step at 3 planes, 1.902 at 4, 1.920 at 8, 1.949 at 16)
    return el3_build_shell(d, in, d.zchunk > 0 ? d.zchunk : kEl3ShellChunk);
}

// CPML-shell tiles (see El3Tile) and their compact memory-variable and CPML
// coefficient storage.  Tiles are listed plane-chunk by plane-chunk, rows
// then columns, so that consecutive tiles share stencil halos.
int el3_build_shell(Device &d, const double *const *in, int kc)
{
    host::ShellTiles st = host::shell_tiles(d.nx, d.ny, d.nz, d.box, d.xoff, d.pitch, kc);
    const std::vector<El3Tile> &tiles = st.tiles;
    d.ntiles = (int)tiles.size();
    d.nslots = st.nslots;
    if (d.ntiles == 0) return SDT_OK;
    if (hipMalloc(&d.tiles, tiles.size() * sizeof(El3Tile)) != hipSuccess) return SDT_ERR_HIP;
    if (hipMemcpy(d.tiles, tiles.data(), tiles.size() * sizeof(El3Tile), hipMemcpyHostToDevice) != hipSuccess)
        return SDT_ERR_HIP;
    // CPML coefficients of every slot: a, b, 1/kappa, a_half, b_half, 1/kappa_half
    // (identity values for slots outside the grid)
    const int cids[6] = {SDT_ACOEF, SDT_BCOEF, SDT_KAPPA, SDT_ACOEF_HALF, SDT_BCOEF_HALF, SDT_KAPPA_HALF};
    auto coef = [&](int q, int i, int j, int k) {
        const double *src = in[cids[q]];
        const bool inverse = (q == 2 || q == 5);
        double v = inverse ? 1.0 : 0.0;
        if (src && i >= 1 && i <= d.nx && j >= 1 && j <= d.ny)
            v = src[(int64_t)(i - 1) + (int64_t)d.nx * ((int64_t)(j - 1) + (int64_t)d.ny * (k - 1))];
        return inverse ? 1.0 / v : v;
    };
    // slot -> cell, for the dictionary's fill
    std::vector<int> cell((size_t)d.nslots * 3);
    host::for_shell_slots(tiles, [&](int64_t slot, int i, int j, int k) {
        cell[(size_t)slot * 3] = i;
        cell[(size_t)slot * 3 + 1] = j;
        cell[(size_t)slot * 3 + 2] = k;
    });
    // Coefficient dictionary: CPML profiles give few distinct tuples (tens
    // for the usual per-axis profiles), so the shell keeps a one-byte id per
    // slot and a small table (rows of 8, 6 used) when there are at most 256.
    const host::RowDictionary cd = host::dedupe_rows(d.nslots, 8, [&](int64_t slot, double *r) {
        const int *c = &cell[(size_t)slot * 3];
        for (int q = 0; q < 6; ++q) r[q] = coef(q, c[0], c[1], c[2]);
    });
    const bool dict = cd.ok;
    d.shell_snq = dict ? M_TOTAL : S_NQ;
    if (dict) {
        if (hipMalloc((void **)&d.shell_cid, (size_t)d.nslots) != hipSuccess) return SDT_ERR_HIP;
        if (hipMemcpy(d.shell_cid, cd.ids.data(), (size_t)d.nslots, hipMemcpyHostToDevice) != hipSuccess)
            return SDT_ERR_HIP;
        if (upload_converted(d, cd.rows.data(), &d.shell_ctab, (int64_t)cd.rows.size()) != SDT_OK) return SDT_ERR_HIP;
        d.shell_ncoef = cd.count();
    }
    // Shell storage in tile-plane blocks: the 256 slots of a tile-plane keep
    // their snq streams together, stream q of slot s at
    // ((s / 256) * snq + q) * 256 + s % 256 (the M_TOTAL memory variables,
    // zero at the start, then without the dictionary the six coefficient
    // streams).
    if (alloc_zeroed(d, &d.shell, (int64_t)d.shell_snq * d.nslots) != SDT_OK) return SDT_ERR_HIP;
    if (!dict) {
        std::vector<double> cs((size_t)d.nslots);
        const size_t row = (size_t)kShellThreads * d.esize, nblk = (size_t)d.nslots / kShellThreads;
        for (int q = 0; q < 6; ++q) {
            for (int64_t slot = 0; slot < d.nslots; ++slot) {
                const int *c = &cell[(size_t)slot * 3];
                cs[(size_t)slot] = coef(q, c[0], c[1], c[2]);
            }
            char *dst = (char *)d.shell + (size_t)(S_CF + q) * row;
            const size_t dpitch = (size_t)d.shell_snq * row;
            hipError_t e;
            if (d.precision == SDT_FP64) {
                e = hipMemcpy2D(dst, dpitch, cs.data(), row, row, nblk, hipMemcpyHostToDevice);
            } else {
                std::vector<float> f(cs.begin(), cs.end());
                e = hipMemcpy2D(dst, dpitch, f.data(), row, row, nblk, hipMemcpyHostToDevice);
            }
            if (e != hipSuccess) return SDT_ERR_HIP;
        }
    }
    return SDT_OK;
}

}  // namespace sdt
