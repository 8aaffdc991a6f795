// host_layout.h -- host-side layout logic of the solver library, free of HIP
// so that it also builds with a plain host compiler (and its sanitizers):
//
//   scan_box        the interior box of CPML-free cells around the grid
//                   centre and the tiles inside it that hold CPML cells
//   el3_interior_box  that box as the 3D elastic interior kernels take it
//   row_layout      padded rows: pitch and first-column offset per grid
//   shell_tiles     the 3D elastic CPML-shell tile list and slot numbering
//   dedupe_rows     up-to-256-entry dictionaries (material tables, shell
//                   CPML coefficients): one-byte ids and distinct rows
//   batch_order     tile lists of a batch in XCD-contiguous order
//   pack_rows /     contiguous padded host images of grid arrays
//   unpack_rows
//
// Grid cells are 1-based (i, j, k) in Fortran order, x fastest, as in the
// reference (cpmlfdtd.f08); host arrays are unpadded nx * ny * nz.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <unordered_map>
#include <vector>

namespace sdt {

// One z-marching tile of the CPML shell: up to kShellThreads cells of a plane,
// planes k0..k1, region kind (0 z slab; 1, 3 the y strip below, above the box;
// 2, 4 the x strip left, right of it) and the first slot of its memory
// variables (slot = base + (k - k0) * kShellThreads + thread).  Over the z
// slabs and y strips (tw = 0) thread t takes element x0 + t of the plane, rows
// of `pitch` elements with column 1 at element xoff of its row: every
// wavefront covers 64 consecutive elements (whole lines) and only the row
// padding leaves threads idle.  Over the x strips (tw = -w for a strip w
// columns wide from column x0, rows from y0) thread t takes the strip's cell
// n0 + t, row by row: column x0 + (n0 + t) mod w of row y0 + (n0 + t) / w, so
// no thread is idle whatever the width.
constexpr int kShellThreads = 256;
struct El3Tile {
    int x0, y0, k0, k1;
    int kind, tw;
    int64_t base;
    int pitch, xoff;
    int n0;
};

// Column (i, j) of thread t of a tile (possibly outside the grid or the
// tile's region: such threads own slots but no cell).  The kernels use the
// same mapping.
inline bool tile_lane(const El3Tile &t, int thread, int &i, int &j)
{
    if (t.tw == 0) {
        const int e = t.x0 + thread;
        j = e / t.pitch + 1;
        i = e - (j - 1) * t.pitch - t.xoff + 1;
        return true;
    }
    const int w = -t.tw, n = t.n0 + thread, r = n / w;   // x strip
    i = t.x0 + (n - r * w);
    j = t.y0 + r;
    return true;
}

namespace host {

// ------------------------------------------------------------ interior box
// A cell is CPML-free when both coefficient sets are the identity (a = 0,
// kappa = 1, b finite); absent arrays read as the identity.
inline bool cpml_free(const double *a, const double *b, const double *k, const double *ah, const double *bh,
                      const double *kh, int64_t c)
{
    auto id = [](double av, double bv, double kv) { return av == 0.0 && kv == 1.0 && std::isfinite(bv); };
    return id(a ? a[c] : 0.0, b ? b[c] : 0.0, k ? k[c] : 1.0) && id(ah ? ah[c] : 0.0, bh ? bh[c] : 0.0, kh ? kh[c] : 1.0);
}

struct BoxScan {
    int box[6];                    // i0, i1, j0, j1, k0, k1 (empty: i1 < i0)
    int tiles_x, tiles_y, tiles_z;
    std::vector<uint8_t> flags;    // tile holds a CPML cell inside the box
    int nflagged;
};

// The box grows from the centre cell along each axis while cells stay
// CPML-free; tiles (tile_x * tile_y * tile_z cells) of the box that still hold
// a CPML cell are flagged.
inline BoxScan scan_box(int nx, int ny, int nz, int tile_x, int tile_y, int tile_z, const double *a, const double *b,
                        const double *k, const double *ah, const double *bh, const double *kh)
{
    BoxScan s;
    s.tiles_x = (nx + tile_x - 1) / tile_x;
    s.tiles_y = (ny + tile_y - 1) / tile_y;
    s.tiles_z = (nz + tile_z - 1) / tile_z;
    s.flags.assign((size_t)s.tiles_x * s.tiles_y * s.tiles_z, 0);
    s.nflagged = 0;
    auto ident = [&](int i, int j, int kk) {
        return cpml_free(a, b, k, ah, bh, kh, (int64_t)(i - 1) + (int64_t)nx * ((int64_t)(j - 1) + (int64_t)ny * (kk - 1)));
    };
    const int ic = (nx + 1) / 2, jc = (ny + 1) / 2, kc = (nz + 1) / 2;
    int *bx = s.box;
    if (!ident(ic, jc, kc)) {
        bx[0] = 1; bx[1] = 0; bx[2] = 1; bx[3] = 0; bx[4] = 1; bx[5] = 0;
        return s;
    }
    bx[0] = ic; while (bx[0] > 1 && ident(bx[0] - 1, jc, kc)) --bx[0];
    bx[1] = ic; while (bx[1] < nx && ident(bx[1] + 1, jc, kc)) ++bx[1];
    bx[2] = jc; while (bx[2] > 1 && ident(ic, bx[2] - 1, kc)) --bx[2];
    bx[3] = jc; while (bx[3] < ny && ident(ic, bx[3] + 1, kc)) ++bx[3];
    bx[4] = kc; while (bx[4] > 1 && ident(ic, jc, bx[4] - 1)) --bx[4];
    bx[5] = kc; while (bx[5] < nz && ident(ic, jc, bx[5] + 1)) ++bx[5];
    for (int kk = bx[4]; kk <= bx[5]; ++kk)
        for (int j = bx[2]; j <= bx[3]; ++j)
            for (int i = bx[0]; i <= bx[1]; ++i)
                if (!ident(i, j, kk)) {
                    const size_t t = (size_t)((i - 1) / tile_x) +
                                     (size_t)s.tiles_x * ((size_t)((j - 1) / tile_y) + (size_t)s.tiles_y * ((kk - 1) / tile_z));
                    s.nflagged += s.flags[t] ? 0 : 1;
                    s.flags[t] = 1;
                }
    return s;
}

// The 3D elastic interior kernels update every cell of the box with all four
// stress nests and all three velocity nests, whose ranges (cpmlfdtd.f08:
// 1761-2127) intersect to [3, n - 2] per axis, and take no CPML cell: the box
// is clipped to that range, and emptied when a tile inside it holds a CPML
// cell (every cell then takes the shell kernels).
inline void el3_interior_box(int (&box)[6], int nflagged, int nx, int ny, int nz)
{
    const int lim[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
        box[2 * a] = box[2 * a] < 3 ? 3 : box[2 * a];
        box[2 * a + 1] = box[2 * a + 1] > lim[a] - 2 ? lim[a] - 2 : box[2 * a + 1];
    }
    const bool empty = nflagged > 0 || box[1] < box[0] || box[3] < box[2] || box[5] < box[4];
    if (empty)
        for (int a = 0; a < 3; ++a) { box[2 * a] = 1; box[2 * a + 1] = 0; }
}

// ------------------------------------------------------------ padded rows
// Row (j, k) holds x cells 1..nx at elements xoff .. xoff + nx - 1 of a row of
// pitch elements: pitch a multiple of the 128-byte line, and cell `first`
// (the box's first column, or 1) at the start of a line.
struct RowLayout {
    int64_t xoff;
    int pitch;
};

inline RowLayout row_layout(int nx, int first, int esize)
{
    const int line = 128 / esize;
    RowLayout r;
    r.xoff = (line - (first - 1) % line) % line;
    r.pitch = (int)((r.xoff + nx + line - 1) / line * line);
    return r;
}

// Contiguous padded image (rows x pitch elements of ncomp values, zero
// padding) of an unpadded host array of rows x nx elements, converted to T.
template <typename T>
void pack_rows(const double *host, size_t rows, size_t nx, size_t pitch, size_t xoff, int ncomp, std::vector<T> &img)
{
    const size_t w = nx * ncomp, p = pitch * ncomp, off = xoff * ncomp;
    img.assign(rows * p, T(0));
    for (size_t r = 0; r < rows; ++r) {
        T *dst = img.data() + r * p + off;
        const double *src = host + r * w;
        for (size_t i = 0; i < w; ++i) dst[i] = (T)src[i];
    }
}

template <typename T>
void unpack_rows(const std::vector<T> &img, size_t rows, size_t nx, size_t pitch, size_t xoff, int ncomp, double *host)
{
    const size_t w = nx * ncomp, p = pitch * ncomp, off = xoff * ncomp;
    for (size_t r = 0; r < rows; ++r) {
        const T *src = img.data() + r * p + off;
        double *dst = host + r * w;
        for (size_t i = 0; i < w; ++i) dst[i] = (double)src[i];
    }
}

// ------------------------------------------------------- 3D elastic shell
// Whether column (i, j) of a tile of region `kind` is a shell cell: z slabs
// take every column; y strips the rows below (kind 1) or above (3) the box's
// y range; x strips the box's y range left (2) or right (4) of its x range.
// Each side keeps to its own cells, so a tile that reaches across a box
// thinner than itself covers nothing twice.  The kernels evaluate the same
// test.
inline bool shell_region(int kind, const int (&b)[6], int i, int j, int nx, int ny)
{
    if (i < 1 || i > nx || j < 1 || j > ny) return false;
    switch (kind) {
    case 0: return true;
    case 1: return j < b[2];
    case 3: return j > b[3];
    case 2: return j >= b[2] && j <= b[3] && i < b[0];
    default: return j >= b[2] && j <= b[3] && i > b[1];
    }
}

struct ShellTiles {
    std::vector<El3Tile> tiles;
    int64_t nslots;
};

// Tiles of the cells outside the box (all cells when the box is empty), plane
// chunk by plane chunk of kc planes: z slabs below and above the box and y
// strips as linear tiles of 256 consecutive elements of the padded rows
// (starting on the 128-byte line grid); x strips (left, then right of the
// box) with their cells numbered row by row, 256 to a tile, so that no lane
// is idle whatever the strip's width.  A tile is listed when one of its
// lanes lies in its region; its slots follow the previous tile's.
inline ShellTiles shell_tiles(int nx, int ny, int nz, const int (&b)[6], int64_t xoff, int pitch, int kc)
{
    ShellTiles st;
    st.nslots = 0;
    const bool empty = b[1] < b[0];
    auto add = [&](int x0, int y0, int k0, int k1, int kind, int tw, int n0 = 0) {
        const El3Tile t{x0, y0, k0, k1, kind, tw, st.nslots, pitch, (int)xoff, n0};
        for (int th = 0; th < kShellThreads; ++th) {
            int i, j;
            if (tile_lane(t, th, i, j) && shell_region(kind, b, i, j, nx, ny)) {
                st.tiles.push_back(t);
                st.nslots += (int64_t)(k1 - k0 + 1) * kShellThreads;
                return;
            }
        }
    };
    // rows ja..jb of planes k0..k1 as linear tiles (elements from row ja's first)
    auto rows = [&](int ja, int jb, int k0, int k1, int kind) {
        for (int e = (ja - 1) * pitch; e < jb * pitch; e += kShellThreads) add(e, 0, k0, k1, kind, 0);
    };
    auto zslab = [&](int ka, int kb) {
        for (int k0 = ka; k0 <= kb; k0 += kc) {
            const int k1 = k0 + kc - 1 < kb ? k0 + kc - 1 : kb;
            rows(1, ny, k0, k1, 0);
        }
    };
    auto xstrip = [&](int lo, int hi, int k0, int k1, int kind) {
        if (lo > hi) return;
        const int w = hi - lo + 1, ncell = w * (b[3] - b[2] + 1);
        for (int n0 = 0; n0 < ncell; n0 += kShellThreads) add(lo, b[2], k0, k1, kind, -w, n0);
    };
    if (empty) {
        zslab(1, nz);
        return st;
    }
    zslab(1, b[4] - 1);
    for (int k0 = b[4]; k0 <= b[5]; k0 += kc) {
        const int k1 = k0 + kc - 1 < b[5] ? k0 + kc - 1 : b[5];
        if (b[2] > 1) rows(1, b[2] - 1, k0, k1, 1);
        xstrip(1, b[0] - 1, k0, k1, 2);
        xstrip(b[1] + 1, nx, k0, k1, 4);
        if (b[3] < ny) rows(b[3] + 1, ny, k0, k1, 3);
    }
    zslab(b[5] + 1, nz);
    return st;
}

// f(slot, i, j, k) for every slot of every tile (threads outside the grid,
// the region or the tile's full rows included, with their unclamped column)
template <typename F>
void for_shell_slots(const std::vector<El3Tile> &tiles, F &&f)
{
    for (const El3Tile &t : tiles)
        for (int k = t.k0; k <= t.k1; ++k)
            for (int th = 0; th < kShellThreads; ++th) {
                int i, j;
                (void)tile_lane(t, th, i, j);
                f(t.base + (int64_t)(k - t.k0) * kShellThreads + th, i, j, k);
            }
}

// ------------------------------------------------------------ dictionaries
// Distinct rows of row_len doubles (compared bit for bit) in order of first
// appearance, and the id of each of n entries; ok = false (and the rest
// unspecified) when there are more than max_rows distinct rows.  fill(e, row)
// writes entry e's row into a zeroed buffer.
struct RowDictionary {
    bool ok;
    std::vector<double> rows;
    std::vector<uint8_t> ids;
    int count() const { return ok && row_len ? (int)(rows.size() / (size_t)row_len) : 0; }
    int row_len;
};

inline RowDictionary dedupe_rows(int64_t n, int row_len, const std::function<void(int64_t, double *)> &fill,
                                 int max_rows = 256)
{
    RowDictionary d;
    d.ok = true;
    d.row_len = row_len;
    const size_t rb = (size_t)row_len * sizeof(double);
    auto hash = [&](const double *r) {
        uint64_t h = 1469598103934665603ull;
        for (int q = 0; q < row_len; ++q) {
            uint64_t w;
            std::memcpy(&w, &r[q], 8);
            h = (h ^ w) * 1099511628211ull;
            h ^= h >> 29;
        }
        return h;
    };
    std::unordered_multimap<uint64_t, int> index;
    d.ids.assign((size_t)n, 0);
    std::vector<double> cur((size_t)row_len), prev((size_t)row_len);
    int prev_id = -1;
    for (int64_t e = 0; e < n; ++e) {
        std::fill(cur.begin(), cur.end(), 0.0);
        fill(e, cur.data());
        if (prev_id >= 0 && std::memcmp(cur.data(), prev.data(), rb) == 0) {   // runs of equal rows
            d.ids[(size_t)e] = (uint8_t)prev_id;
            continue;
        }
        const uint64_t h = hash(cur.data());
        int id = -1;
        auto range = index.equal_range(h);
        for (auto it = range.first; it != range.second; ++it)
            if (std::memcmp(&d.rows[(size_t)it->second * row_len], cur.data(), rb) == 0) { id = it->second; break; }
        if (id < 0) {
            if ((int)(d.rows.size() / row_len) == max_rows) {
                d.ok = false;
                return d;
            }
            id = (int)(d.rows.size() / row_len);
            d.rows.insert(d.rows.end(), cur.begin(), cur.end());
            index.emplace(h, id);
        }
        d.ids[(size_t)e] = (uint8_t)id;
        prev.swap(cur);
        prev_id = id;
    }
    return d;
}

// ---------------------------------------------------------------- batches
// One tile of a batched 2D launch: shot s (-1: padding), tile column x, tile
// row z; 16 bytes, the layout of the device's int4.
struct BatchTile {
    int s, x, z, w;
};

// XCD-contiguous order of a tile list: natural tile q belongs to group
// g = q / 8, which the hardware's round-robin dispatch sends to XCD g % 8;
// the list is padded to a multiple of 64 with shot -1 and tile q moves to
// position ((g / 8) * 8 + q % 8) * 8 + g % 8, so each XCD receives runs of
// neighbouring tiles.
inline std::vector<BatchTile> batch_order(const std::vector<BatchTile> &t)
{
    std::vector<BatchTile> in(t);
    in.resize((in.size() + 63) / 64 * 64, BatchTile{-1, 0, 0, 0});
    std::vector<BatchTile> p(in.size());
    for (size_t q = 0; q < in.size(); ++q) {
        const size_t g = q / 8, slot = (g / 8) * 8 + q % 8;
        p[slot * 8 + g % 8] = in[q];
    }
    return p;
}

}  // namespace host
}  // namespace sdt
