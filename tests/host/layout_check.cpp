// layout_check.cpp -- host-side layout logic of the solver library
// (seidart-backend_amd/csrc/host_layout.h) under AddressSanitizer and
// UndefinedBehaviorSanitizer, over seeded random grids:
//
//   * scan_box / el3_interior_box: the box holds only CPML-free cells, lies in
//     [3, n - 2] per axis, and is empty when a CPML cell sits inside it;
//   * row_layout: the box's first column starts a 128-byte line, the pitch is
//     whole lines and holds the row;
//   * shell_tiles: every cell outside the box is covered by exactly one
//     (tile, lane, plane), no cell inside it is, z-slab and y-strip tiles
//     start on the line grid and x-strip tiles inside their strip, chunks hold
//     at most kc planes, and slots number the tiles' lanes without gaps or
//     overlaps;
//   * dedupe_rows: every entry's id maps back to its row bit for bit, ids are
//     numbered in order of first appearance, and more than 256 distinct rows
//     is reported;
//   * pack_rows / unpack_rows: round trips, zero padding;
//   * batch_order: a permutation of the tiles plus padding that sends group g
//     of 8 consecutive tiles to XCD g % 8.
//
// Exits non-zero (or the sanitizers abort) on the first failure.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../seidart-backend_amd/csrc/host_layout.h"

using namespace sdt;

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static uint64_t next_u64()
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return rng_state;
}
static int uni(int lo, int hi) { return lo + (int)(next_u64() % (uint64_t)(hi - lo + 1)); }
static double unif() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

static int failures = 0;
#define CHECK(cond, ...)                                            \
    do {                                                            \
        if (!(cond)) {                                              \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);                      \
            std::fprintf(stderr, "\n");                             \
            if (++failures > 20) std::exit(1);                      \
        }                                                           \
    } while (0)

// CPML arrays of a grid with a w-cell absorbing layer on every face (a != 0,
// kappa > 1 there), optionally with stray CPML cells inside
struct Cpml {
    std::vector<double> a, b, k, ah, bh, kh;
};

static Cpml make_cpml(int nx, int ny, int nz, int w, int strays)
{
    const size_t n = (size_t)nx * ny * nz;
    Cpml c;
    c.a.assign(n, 0.0); c.b.assign(n, 0.0); c.k.assign(n, 1.0);
    c.ah.assign(n, 0.0); c.bh.assign(n, 0.0); c.kh.assign(n, 1.0);
    for (int kk = 1; kk <= nz; ++kk)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const int d = std::min(std::min(std::min(i, nx + 1 - i), std::min(j, ny + 1 - j)),
                                       std::min(kk, nz + 1 - kk));
                if (d <= w) {
                    const size_t e = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (kk - 1));
                    const double s = (double)(w + 1 - d) / (w + 1);
                    c.a[e] = -0.1 * s; c.b[e] = 1.0 - 0.2 * s; c.k[e] = 1.0 + s;
                    c.ah[e] = -0.1 * s; c.bh[e] = 1.0 - 0.2 * s; c.kh[e] = 1.0 + s;
                }
            }
    for (int q = 0; q < strays; ++q) {
        const size_t e = (size_t)(next_u64() % n);
        c.kh[e] = 1.5;
    }
    return c;
}

static void check_grid(int nx, int ny, int nz, int w, int strays, int esize, int kc)
{
    const Cpml c = make_cpml(nx, ny, nz, w, strays);
    host::BoxScan bs = host::scan_box(nx, ny, nz, 64, 4, 1, c.a.data(), c.b.data(), c.k.data(), c.ah.data(),
                                      c.bh.data(), c.kh.data());
    CHECK(bs.flags.size() == (size_t)bs.tiles_x * bs.tiles_y * bs.tiles_z, "flag count");
    int box[6];
    for (int q = 0; q < 6; ++q) box[q] = bs.box[q];
    host::el3_interior_box(box, bs.nflagged, nx, ny, nz);
    const bool empty = box[1] < box[0];
    auto at = [&](int i, int j, int kk) {
        return (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (kk - 1));
    };
    if (!empty) {
        CHECK(box[0] >= 3 && box[1] <= nx - 2 && box[2] >= 3 && box[3] <= ny - 2 && box[4] >= 3 && box[5] <= nz - 2,
              "box outside [3, n-2]: %d..%d %d..%d %d..%d of %dx%dx%d", box[0], box[1], box[2], box[3], box[4],
              box[5], nx, ny, nz);
        for (int kk = box[4]; kk <= box[5]; ++kk)
            for (int j = box[2]; j <= box[3]; ++j)
                for (int i = box[0]; i <= box[1]; ++i)
                    CHECK(host::cpml_free(c.a.data(), c.b.data(), c.k.data(), c.ah.data(), c.bh.data(), c.kh.data(),
                                          (int64_t)at(i, j, kk)),
                          "CPML cell (%d,%d,%d) inside the box", i, j, kk);
    } else {
        CHECK(box[0] == 1 && box[1] == 0 && box[2] == 1 && box[3] == 0 && box[4] == 1 && box[5] == 0, "empty box form");
    }
    if (strays == 0 && w >= 2 && nx > 2 * w + 4 && ny > 2 * w + 4 && nz > 2 * w + 4)
        CHECK(!empty && box[0] == w + 1 && box[1] == nx - w, "box of a %d-cell layer: %d..%d", w, box[0], box[1]);

    const int line = 128 / esize;
    const host::RowLayout rl = host::row_layout(nx, empty ? 1 : box[0], esize);
    CHECK(rl.xoff >= 0 && rl.xoff < line, "xoff %lld", (long long)rl.xoff);
    CHECK((rl.xoff + (empty ? 1 : box[0]) - 1) % line == 0, "first column off the line grid");
    CHECK(rl.pitch % line == 0 && rl.pitch >= rl.xoff + nx && rl.pitch < rl.xoff + nx + line, "pitch %d", rl.pitch);

    const host::ShellTiles st = host::shell_tiles(nx, ny, nz, box, rl.xoff, rl.pitch, kc);
    std::vector<int> cover((size_t)nx * ny * nz, 0);
    int64_t base = 0;
    for (const El3Tile &t : st.tiles) {
        CHECK(t.base == base, "tile base %lld, expected %lld", (long long)t.base, (long long)base);
        CHECK(t.k0 >= 1 && t.k1 <= nz && t.k0 <= t.k1 && t.k1 - t.k0 + 1 <= kc, "planes %d..%d", t.k0, t.k1);
        const bool xs = t.kind == 2 || t.kind == 4;
        CHECK(t.kind >= 0 && t.kind <= 4 && (xs ? t.tw < 0 : t.tw == 0), "tile shape of kind %d", t.kind);
        // linear tiles (z slabs, y strips) start on a line, inside the plane
        CHECK(xs || (t.x0 % line == 0 && t.x0 >= 0 && t.x0 < ny * rl.pitch && t.pitch == rl.pitch &&
                     t.xoff == rl.xoff),
              "linear tile at element %d", t.x0);
        if (xs)   // x-strip tiles cover their whole strip's width from its first row
            CHECK((t.kind == 2 ? t.x0 == 1 && -t.tw == box[0] - 1 : t.x0 == box[1] + 1 && -t.tw == nx - box[1]) &&
                      t.y0 == box[2] && t.n0 >= 0 && t.n0 % kShellThreads == 0,
                  "x-strip tile x0 %d width %d", t.x0, -t.tw);
        bool any = false;
        for (int th = 0; th < kShellThreads; ++th) {
            int i, j;
            if (!tile_lane(t, th, i, j) || !host::shell_region(t.kind, box, i, j, nx, ny)) continue;
            any = true;
            for (int kk = t.k0; kk <= t.k1; ++kk) {
                const bool inside = !empty && i >= box[0] && i <= box[1] && j >= box[2] && j <= box[3] &&
                                    kk >= box[4] && kk <= box[5];
                CHECK(!inside, "box cell (%d,%d,%d) in a shell tile", i, j, kk);
                ++cover[at(i, j, kk)];
            }
        }
        CHECK(any, "tile without a shell cell");
        base += (int64_t)(t.k1 - t.k0 + 1) * kShellThreads;
    }
    CHECK(st.nslots == base, "nslots %lld, tiles hold %lld", (long long)st.nslots, (long long)base);
    for (int kk = 1; kk <= nz; ++kk)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const bool inside = !empty && i >= box[0] && i <= box[1] && j >= box[2] && j <= box[3] &&
                                    kk >= box[4] && kk <= box[5];
                CHECK(cover[at(i, j, kk)] == (inside ? 0 : 1), "cell (%d,%d,%d) covered %d times (%dx%dx%d w%d kc%d)",
                      i, j, kk, cover[at(i, j, kk)], nx, ny, nz, w, kc);
            }
    std::vector<uint8_t> seen((size_t)st.nslots, 0);
    host::for_shell_slots(st.tiles, [&](int64_t slot, int, int, int kk) {
        CHECK(slot >= 0 && slot < st.nslots, "slot %lld", (long long)slot);
        if (slot >= 0 && slot < st.nslots) {
            CHECK(!seen[(size_t)slot], "slot %lld twice", (long long)slot);
            seen[(size_t)slot] = 1;
        }
        CHECK(kk >= 1 && kk <= nz, "slot plane %d", kk);
    });
    for (size_t s = 0; s < seen.size(); ++s) CHECK(seen[s], "slot %zu not visited", s);
}

static void check_dictionary(int64_t n, int row_len, int distinct)
{
    std::vector<std::vector<double>> pool((size_t)distinct, std::vector<double>((size_t)row_len));
    for (auto &r : pool)
        for (double &v : r) v = (double)uni(0, 3) * 0.5;
    for (int q = 0; q < distinct; ++q) pool[(size_t)q][0] = (double)q;   // rows differ
    if (distinct > 1) pool[1][0] = -0.0, pool[0][0] = 0.0;               // -0 and +0 are different bit patterns
    std::vector<int> pick((size_t)n);
    for (auto &p : pick) p = next_u64() % 4 == 0 ? uni(0, distinct - 1) : (int)(&p - pick.data()) % distinct;
    auto fill = [&](int64_t e, double *r) {
        for (int q = 0; q < row_len; ++q) r[q] = pool[(size_t)pick[(size_t)e]][(size_t)q];
    };
    std::vector<uint8_t> used((size_t)distinct, 0);
    int used_count = 0;
    for (int p : pick) used_count += used[(size_t)p] ? 0 : (used[(size_t)p] = 1);
    const host::RowDictionary d = host::dedupe_rows(n, row_len, fill);
    CHECK(d.ok == (used_count <= 256), "ok %d with %d distinct rows", (int)d.ok, used_count);
    if (!d.ok) return;
    CHECK(d.count() == used_count, "count %d, expected %d", d.count(), used_count);
    std::vector<double> row((size_t)row_len);
    int maxid = -1;
    for (int64_t e = 0; e < n; ++e) {
        const int id = d.ids[(size_t)e];
        CHECK(id < d.count(), "id %d of %d", id, d.count());
        CHECK(id <= maxid + 1, "id %d before %d", id, maxid + 1);
        maxid = std::max(maxid, id);
        std::fill(row.begin(), row.end(), 0.0);
        fill(e, row.data());
        CHECK(std::memcmp(&d.rows[(size_t)id * row_len], row.data(), sizeof(double) * row_len) == 0,
              "entry %lld maps to another row", (long long)e);
    }
}

template <typename T>
static void check_image(size_t rows, size_t nx, int esize, int ncomp, int first)
{
    const host::RowLayout rl = host::row_layout((int)nx, first, esize);
    std::vector<double> src(rows * nx * ncomp), back(rows * nx * ncomp, -1.0);
    for (double &v : src) v = unif() * 2.0 - 1.0;
    std::vector<T> img;
    host::pack_rows<T>(src.data(), rows, nx, (size_t)rl.pitch, (size_t)rl.xoff, ncomp, img);
    CHECK(img.size() == rows * rl.pitch * ncomp, "image size");
    for (size_t r = 0; r < rows; ++r)
        for (size_t e = 0; e < (size_t)rl.pitch * ncomp; ++e) {
            const bool pad = e < (size_t)rl.xoff * ncomp || e >= ((size_t)rl.xoff + nx) * ncomp;
            if (pad) CHECK(img[r * rl.pitch * ncomp + e] == T(0), "padding not zero");
        }
    host::unpack_rows<T>(img, rows, nx, (size_t)rl.pitch, (size_t)rl.xoff, ncomp, back.data());
    for (size_t e = 0; e < src.size(); ++e) CHECK(back[e] == (double)(T)src[e], "round trip at %zu", e);
}

static void check_batch(int ntiles)
{
    std::vector<host::BatchTile> t((size_t)ntiles);
    for (int q = 0; q < ntiles; ++q) t[(size_t)q] = host::BatchTile{q, q % 7, q / 7, 0};
    const std::vector<host::BatchTile> p = host::batch_order(t);
    CHECK(p.size() % 64 == 0 && p.size() >= (size_t)ntiles && p.size() < (size_t)ntiles + 64, "padded size %zu",
          p.size());
    std::vector<int> seen((size_t)ntiles, 0);
    size_t pads = 0;
    for (size_t pos = 0; pos < p.size(); ++pos) {
        const int q = p[pos].s;
        if (q < 0) { ++pads; continue; }
        CHECK(q < ntiles, "tile %d", q);
        if (q >= ntiles) continue;
        ++seen[(size_t)q];
        CHECK((size_t)((q / 8) % 8) == pos % 8, "tile %d at position %zu: XCD %zu, group %d", q, pos, pos % 8, q / 8);
    }
    for (int q = 0; q < ntiles; ++q) CHECK(seen[(size_t)q] == 1, "tile %d seen %d times", q, seen[(size_t)q]);
    CHECK(pads == p.size() - (size_t)ntiles, "padding count");
}

int main(int argc, char **argv)
{
    const int rounds = argc > 1 ? std::atoi(argv[1]) : 300;
    // the 3D elastic shapes the tests and the benchmark use, then random ones
    const int fixed[][7] = {{40, 36, 32, 6, 0, 8, 8}, {28, 28, 28, 6, 0, 8, 8}, {32, 30, 28, 6, 0, 4, 8},
                            {20, 18, 16, 4, 3, 8, 3}, {7, 7, 7, 2, 0, 8, 1}, {5, 5, 5, 0, 0, 4, 8},
                            {1, 1, 1, 0, 0, 8, 8}, {70, 9, 40, 10, 0, 8, 12},
                            // CPML strips wider than a wavefront (66 cells: a strip row spans two)
                            {140, 140, 140, 66, 0, 8, 3}, {141, 139, 138, 66, 0, 4, 3}};
    for (const auto &f : fixed) check_grid(f[0], f[1], f[2], f[3], f[4], f[5], f[6]);
    for (int r = 0; r < rounds; ++r) {
        const int nx = uni(1, 90), ny = uni(1, 40), nz = uni(1, 30);
        check_grid(nx, ny, nz, uni(0, 12), next_u64() % 5 == 0 ? uni(1, 3) : 0, next_u64() % 2 ? 8 : 4, uni(1, 12));
    }
    for (int r = 0; r < 40; ++r) {
        const int distinct = (r % 4 == 0) ? uni(257, 320) : uni(1, 256);
        check_dictionary(uni(0, 4000), r % 3 == 0 ? 8 : (r % 3 == 1 ? 32 : 1), distinct);
    }
    check_dictionary(0, 8, 1);
    for (int r = 0; r < 30; ++r) {
        const size_t rows = (size_t)uni(1, 50), nx = (size_t)uni(1, 70);
        const int first = uni(1, (int)nx);
        check_image<double>(rows, nx, 8, r % 2 ? 2 : 1, first);
        check_image<float>(rows, nx, 4, r % 2 ? 2 : 1, first);
    }
    for (int n : {0, 1, 7, 8, 63, 64, 65, 200, 1000, 4097}) check_batch(n);
    if (failures) {
        std::fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    std::printf("layout_check: all checks passed\n");
    return 0;
}
