// em3.hip -- 3D anisotropic electromagnetic CPML-FDTD (electromag3,
// cpmlfdtd.f08:3353-4001) on gfx950: (Hx, Hy, Hz) from E, then E from H with
// the full symmetric inverse permittivity and the conductivity tensor applied
// to E_old.  2nd-order Yee differences.
//
// As in em2.hip: E_old == E at each step start, so the E update runs in
// place; the H faces (3921-3961) are zeroed lazily (previous H taken as 0 on a
// face, fresh value kept for this step's E update, masked on read-back), and
// H cells outside every H nest lie on faces and are written 0 directly.
#include <math.h>

#include <vector>

#include "sdt_device.h"
#include "sdt_internal.h"

namespace sdt {

template <typename T>
struct Em3Args {
    T *E[3], *H[3];
    const T *aEx, *bEx, *cEx, *bEy, *cEy, *cEz;            // aEy=bEx, aEz=cEx, bEz=cEy
    const T *s11, *s12, *s13, *s22, *s23, *s33;
    const T *a, *b, *ik, *ah, *bh, *ikh;
    // CPML coefficients as a dictionary (ctab != nullptr): id per padded cell,
    // rows of 8 = a b 1/kappa a_half b_half 1/kappa_half
    const uint8_t *cid;
    const T *ctab;
    // E-side (H update) and H-side (E update) memory variables
    T *m_dEy_dz, *m_dEz_dy, *m_dEx_dz, *m_dEz_dx, *m_dEx_dy, *m_dEy_dx;
    T *m_dHz_dy, *m_dHy_dz, *m_dHz_dx, *m_dHx_dz, *m_dHx_dy, *m_dHy_dx;
    const uint8_t *flags;
    int any_flags;      // 0: no tile inside the box holds a CPML cell (the flags need no load)
    int col_order;      // half-step kernels: 1D grid in XCD column order
    int lin;            // half-step kernels: each workgroup takes 256 consecutive elements of a plane
    int box[6];
    int nx, ny, nz, tiles_x, tiles_y, pitch;
    int64_t xoff;
    T rdx, rdy, rdz, dt, dbH;
    const int *d_it;
    unsigned long long *vnorm;
    int64_t src_cell;
    PwFaces pwf;        // plane-wave face cells, excluded from the fused norm
    // material table (at most 256 distinct materials): id per padded cell,
    // rows of kEmRow = aEx bEx cEx bEy cEy cEz s11 s12 s13 s22 s23 s33
    const uint8_t *mid;
    const T *tab;
    // fused step: the previous step's E, H and H-update memory variables
    // (the new ones go to E, H and m_dEy_dz .. m_dEy_dx)
    const T *Eo[3], *Ho[3];
    const T *mo[6];   // m_dEy_dz m_dEz_dy m_dEx_dz m_dEz_dx m_dEx_dy m_dEy_dx
};

constexpr int kEmRow = 16;
// planes per workgroup of the fused kernel when the z-chunk option is 0 (auto)
constexpr int kEm3FusedChunk = 12;

// A CPML cell's coefficients: from the dictionary when there is one (a
// wave-uniform test of a kernel argument), else from the six arrays.
template <typename T>
__device__ __forceinline__ void em3_coefs(const Em3Args<T> &A, int64_t c, T &a, T &b, T &ik, T &ah, T &bh, T &ikh)
{
    if (A.ctab) {
        const T *r = A.ctab + 8 * (int)A.cid[c];
        a = r[0]; b = r[1]; ik = r[2]; ah = r[3]; bh = r[4]; ikh = r[5];
    } else {
        a = A.a[c]; b = A.b[c]; ik = A.ik[c];
        ah = A.ah[c]; bh = A.bh[c]; ikh = A.ikh[c];
    }
}

template <typename T>
using em_const_row = __attribute__((address_space(4))) const T *;

// Material accessors for the E update: entry q of the row above.
template <typename T>
struct EmScalarRow {   // wave-uniform row: scalar loads through the constant address space
    em_const_row<T> p;
    __device__ T operator[](int q) const { return p[q]; }
};
template <typename T>
struct EmLaneRow {     // per-lane row from the (cached) table
    const T *p;
    __device__ T operator[](int q) const { return p[q]; }
};
template <typename T>
struct EmDenseRow {    // per-cell arrays
    const Em3Args<T> *A;
    int64_t c;
    __device__ T operator[](int q) const
    {
        const T *const arr[12] = {A->aEx, A->bEx, A->cEx, A->bEy, A->cEy, A->cEz,
                                  A->s11, A->s12, A->s13, A->s22, A->s23, A->s33};
        return arr[q][c];
    }
};

// E += dt * inv(eps) * (curl H - sigma E_old), cpmlfdtd.f08:3819-3828
template <typename T, typename R>
__device__ __forceinline__ void em3_e_update(const R &M, T exo, T eyo, T ezo, T cx, T cy, T cz, T dt, T &ex, T &ey,
                                             T &ez)
{
    const T rx = cx - (M[6] * exo + M[7] * eyo + M[8] * ezo);
    const T ry = cy - (M[7] * exo + M[9] * eyo + M[10] * ezo);
    const T rz = cz - (M[8] * exo + M[10] * eyo + M[11] * ezo);
    ex = exo + (M[0] * rx + M[1] * ry + M[2] * rz) * dt;
    ey = eyo + (M[1] * rx + M[3] * ry + M[4] * rz) * dt;
    ez = ezo + (M[2] * rx + M[4] * ry + M[5] * rz) * dt;
}

// ---- cell updates shared by the half-step and fused kernels ----------------
// The kernels read the previous step's E, H and H-update memory variables
// through Eo, Ho, mo and write the new ones through E, H and m_dE*: the same
// arrays in the half-step kernels (updates in place), the other buffers of a
// pair in the fused kernel.
template <typename T>
__device__ __forceinline__ double fmax_nan3(double a, double b) { return (b > a || b != b) ? b : a; }

// H of a cell (cpmlfdtd.f08:3700-3778) from its operands: this step's E of
// the cell and of its +x, +y, +z neighbours and the previous H (0 on the
// lazily zeroed faces).  In CPML cells the coefficients and memory variables
// are read here.  `store` writes H and the memory variables (the cell's
// owner); otherwise they are only returned (a halo cell of the fused kernel).
template <typename T>
__device__ __forceinline__ void em3_h_math(const Em3Args<T> &A, int64_t c, int k, bool pml, bool store, bool mst,
                                           T ex0, T ey0, T ez0, T exy, T ezy, T eyx, T ezx, T exz, T eyz, T h0,
                                           T h1, T h2, T &hx, T &hy, T &hz)
{
    T a = 0, b = 0, ik = 1, ah = 0, bh = 0, ikh = 1;
    T m_eydz = 0, m_ezdy = 0, m_exdz = 0, m_ezdx = 0, m_exdy = 0, m_eydx = 0;
    if (pml) {
        em3_coefs(A, c, a, b, ik, ah, bh, ikh);
        m_eydz = A.mo[0][c]; m_ezdy = A.mo[1][c]; m_exdz = A.mo[2][c];
        m_ezdx = A.mo[3][c]; m_exdy = A.mo[4][c]; m_eydx = A.mo[5][c];
    }
    T dEz_dy = (ezy - ez0) * A.rdy, dEy_dz = (eyz - ey0) * A.rdz;
    T dEx_dz = (exz - ex0) * A.rdz, dEz_dx = (ezx - ez0) * A.rdx;
    T dEx_dy = (exy - ex0) * A.rdy, dEy_dx = (eyx - ey0) * A.rdx;
    if (pml) {
        m_eydz = bh * m_eydz + ah * dEy_dz;
        m_ezdy = bh * m_ezdy + ah * dEz_dy;
        dEz_dy = dEz_dy * ikh + m_ezdy;
        dEy_dz = dEy_dz * ikh + m_eydz;
        m_exdz = b * m_exdz + a * dEx_dz;
        m_ezdx = b * m_ezdx + a * dEz_dx;
        dEx_dz = dEx_dz * ik + m_exdz;
        dEz_dx = dEz_dx * ik + m_ezdx;
        m_exdy = b * m_exdy + a * dEx_dy;
        m_eydx = b * m_eydx + a * dEy_dx;
        dEx_dy = dEx_dy * ik + m_exdy;
        dEy_dx = dEy_dx * ik + m_eydx;
    }
    hx = h0 + A.dbH * (dEy_dz - dEz_dy);
    hy = h1 + A.dbH * (dEz_dx - dEx_dz);
    hz = k >= 2 ? h2 + A.dbH * (dEx_dy - dEy_dx) : T(0);   // Hz nest from k = 2
    if (store) {
        A.H[0][c] = hx; A.H[1][c] = hy; A.H[2][c] = hz;
        if (mst) {   // 0 from the lanes without CPML (their memory variables stay 0)
            A.m_dEy_dz[c] = m_eydz; A.m_dEz_dy[c] = m_ezdy;
            A.m_dEx_dz[c] = m_exdz; A.m_dEz_dx[c] = m_ezdx;
            if (k >= 2) { A.m_dEx_dy[c] = m_exdy; A.m_dEy_dx[c] = m_eydx; }
        }
    }
}

// H of cell (i, j, k) after this step's H update, its operands loaded here
// (the fused kernel).
template <typename T>
__device__ __forceinline__ void em3_h_new(const Em3Args<T> &A, int i, int j, int k, bool store, T &hx, T &hy, T &hz)
{
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int64_t sy = A.pitch, sz = (int64_t)A.pitch * ny;
    const int64_t c = A.xoff + (int64_t)(i - 1) + sy * (j - 1) + sz * (k - 1);
    if (i == nx || j == ny || k == nz) {
        hx = hy = hz = T(0);
        if (store) { A.H[0][c] = T(0); A.H[1][c] = T(0); A.H[2][c] = T(0); }
        return;
    }
    const bool face = i == 1 || j == 1 || k == 1;
    const T *Ex = A.Eo[0], *Ey = A.Eo[1], *Ez = A.Eo[2];
    const T ex0 = Ex[c], exy = Ex[c + sy], exz = Ex[c + sz];
    const T ey0 = Ey[c], eyx = Ey[c + 1], eyz = Ey[c + sz];
    const T ez0 = Ez[c], ezx = Ez[c + 1], ezy = Ez[c + sy];
    const T h0 = face ? T(0) : A.Ho[0][c], h1 = face ? T(0) : A.Ho[1][c], h2 = face ? T(0) : A.Ho[2][c];
    const size_t t = (size_t)((i - 1) >> 6) + (size_t)A.tiles_x * ((size_t)((j - 1) >> 2) + (size_t)A.tiles_y * (k - 1));
    const bool pml = !in_box(A.box, i, j, k) || (A.any_flags && A.flags[t]);
    em3_h_math(A, c, k, pml, store, pml, ex0, ey0, ez0, exy, ezy, eyx, ezx, exz, eyz, h0, h1, h2, hx, hy, hz);
}

// E-update operands of a cell that do not depend on this step's H: the old
// E, the material id and, in CPML cells, the coefficients and memory
// variables.  Loaded before the plane's barrier so their latency overlaps the
// H update.
template <typename T>
struct EPre {
    T exo, eyo, ezo;
    int m;
    bool pml;
    T a, b, ik, ah, bh, ikh;
    T mhzdy, mhydz, mhzdx, mhxdz, mhxdy, mhydx;
};

template <typename T, bool TAB>
__device__ __forceinline__ EPre<T> em3_e_pre(const Em3Args<T> &A, int64_t c, bool pml)
{
    EPre<T> P;
    P.exo = A.Eo[0][c]; P.eyo = A.Eo[1][c]; P.ezo = A.Eo[2][c];
    P.m = TAB ? (int)A.mid[c] : 0;
    P.pml = pml;
    P.a = 0; P.b = 0; P.ik = 1; P.ah = 0; P.bh = 0; P.ikh = 1;
    P.mhzdy = P.mhydz = P.mhzdx = P.mhxdz = P.mhxdy = P.mhydx = T(0);
    if (pml) {
        em3_coefs(A, c, P.a, P.b, P.ik, P.ah, P.bh, P.ikh);
        P.mhzdy = A.m_dHz_dy[c]; P.mhydz = A.m_dHy_dz[c]; P.mhzdx = A.m_dHz_dx[c];
        P.mhxdz = A.m_dHx_dz[c]; P.mhxdy = A.m_dHx_dy[c]; P.mhydx = A.m_dHy_dx[c];
    }
    return P;
}

// E of cell (i, j, k) from this step's H (own cell; x- and y-neighbours; the
// plane below) and the preloaded operands; returns the cell's squared |E| for the norm.
// Columns past nx (the row padding) take 0 like the faces.  The memory
// variables are stored by the CPML cells (WL = false) or, with WL, by every
// lane with mst set (0 from the lanes without CPML, whose memory variables
// stay 0), so that a line holding CPML cells is written whole.
template <typename T, bool TAB, bool WL = false>
__device__ __forceinline__ double em3_e_new(const Em3Args<T> &A, int i, int j, int k, const EPre<T> &P, T hx0,
                                            T hy0, T hz0, T hxy, T hyx, T hzx, T hzy, T hxz, T hyz, bool mst = false)
{
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int64_t sy = A.pitch, sz = (int64_t)A.pitch * ny;
    const int64_t c = A.xoff + (int64_t)(i - 1) + sy * (j - 1) + sz * (k - 1);
    T ex, ey, ez;
    T m_hzdy = 0, m_hydz = 0, m_hzdx = 0, m_hxdz = 0, m_hxdy = 0, m_hydx = 0;
    if (i == 1 || i >= nx || j == 1 || j == ny || k == 1 || k == nz) {
        ex = ey = ez = T(0);
    } else {
        T dHz_dy = (hz0 - hzy) * A.rdy;
        T dHy_dz = (hy0 - hyz) * A.rdz;
        T dHz_dx = (hz0 - hzx) * A.rdx;
        T dHx_dz = (hx0 - hxz) * A.rdz;
        T dHy_dx = (hy0 - hyx) * A.rdx;
        T dHx_dy = (hx0 - hxy) * A.rdy;
        if (P.pml) {
            m_hzdy = P.bh * P.mhzdy + P.ah * dHz_dy;
            m_hydz = P.b * P.mhydz + P.a * dHy_dz;
            m_hzdx = P.bh * P.mhzdx + P.ah * dHz_dx;
            m_hxdz = P.bh * P.mhxdz + P.ah * dHx_dz;
            m_hxdy = P.bh * P.mhxdy + P.ah * dHx_dy;
            m_hydx = P.bh * P.mhydx + P.ah * dHy_dx;
            dHz_dy = dHz_dy * P.ikh + m_hzdy;
            dHy_dz = dHy_dz * P.ik + m_hydz;
            dHz_dx = dHz_dx * P.ikh + m_hzdx;
            dHx_dz = dHx_dz * P.ikh + m_hxdz;
            dHx_dy = dHx_dy * P.ikh + m_hxdy;
            dHy_dx = dHy_dx * P.ikh + m_hydx;
            if (!WL) {
                A.m_dHz_dy[c] = m_hzdy; A.m_dHy_dz[c] = m_hydz; A.m_dHz_dx[c] = m_hzdx;
                A.m_dHx_dz[c] = m_hxdz; A.m_dHx_dy[c] = m_hxdy; A.m_dHy_dx[c] = m_hydx;
            }
        }
        const T cx = dHz_dy - dHy_dz, cy = dHx_dz - dHz_dx, cz = dHy_dx - dHx_dy;
        if constexpr (TAB) {
            const int m0 = __builtin_amdgcn_readfirstlane(P.m);
            if (__all(P.m == m0)) {
                const EmScalarRow<T> M{(em_const_row<T>)(A.tab) + m0 * kEmRow};
                em3_e_update(M, P.exo, P.eyo, P.ezo, cx, cy, cz, A.dt, ex, ey, ez);
            } else {
                const EmLaneRow<T> M{A.tab + P.m * kEmRow};
                em3_e_update(M, P.exo, P.eyo, P.ezo, cx, cy, cz, A.dt, ex, ey, ez);
            }
        } else {
            const EmDenseRow<T> M{&A, c};
            em3_e_update(M, P.exo, P.eyo, P.ezo, cx, cy, cz, A.dt, ex, ey, ez);
        }
    }
    A.E[0][c] = ex; A.E[1][c] = ey; A.E[2][c] = ez;
    if (WL && mst) {
        A.m_dHz_dy[c] = m_hzdy; A.m_dHy_dz[c] = m_hydz; A.m_dHz_dx[c] = m_hzdx;
        A.m_dHx_dz[c] = m_hxdz; A.m_dHx_dy[c] = m_hxdy; A.m_dHy_dx[c] = m_hydx;
    }
    if (c == A.src_cell || pw_face(A.pwf, i, j, k)) return 0.0;
    const double x = ex, y = ey, z = ez;
    return x * x + y * y + z * z;
}

// ---- half-step kernels ----------------------------------------------------
// One kernel for the H update and one for the E update, a 64 x 4 tile of one
// plane per workgroup.  Every operand is loaded before the first store: the
// compiler cannot tell the arrays apart, so a store ahead of a load would
// serialise the memory round trips.
//
// Block order.  The dispatcher deals workgroup b to XCD b % 8, and each XCD
// has its own L2.  In the grid's own order a tile and the tiles above and
// below it (whose E or H planes it reads) land on different XCDs.  In column
// order (a 1D grid) XCD g takes the tile columns m = g, g + 8, g + 16, ... in
// turn and runs each through every plane, so those tiles share one L2 at
// about the same time.
struct Em3Block {
    int bx, by, bz;
    bool valid;
};

template <typename T>
__device__ __forceinline__ Em3Block em3_block(const Em3Args<T> &A)
{
    if (!A.col_order) return {(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, true};
    const unsigned b = blockIdx.x, g = b & 7u, s = b >> 3;
    const unsigned m = (s / (unsigned)A.nz) * 8u + g, kz = s % (unsigned)A.nz;
    const unsigned ncol = (unsigned)A.tiles_x * (unsigned)A.tiles_y;
    return {(int)(m % (unsigned)A.tiles_x), (int)(m / (unsigned)A.tiles_x), (int)kz, m < ncol};
}

// Whole-line stores (line_any): in the half-step kernels each wavefront
// holds one row segment starting on a line, so the lanes over the row padding
// (columns past nx) store zeros with the row, and when a line holds a CPML
// cell every lane of it stores the memory variables (0 from the lanes without
// CPML: their memory variables stay 0, as cpmlfdtd.f08's do where a = 0).

// The cell of a half-step kernel's thread.  With `lin` a workgroup takes 256
// consecutive elements of plane blockIdx.y + 1, rows following one another,
// so that no lane is idle beyond the row padding (64 x 4 tiles of 404-cell
// rows leave 32 of every 448 lanes idle); otherwise a 64 x 4 tile, in grid or
// XCD column order.  `row`: the lane holds a cell or a padding element of a
// row of the grid.
struct Em3Lane {
    int i, j, k;
    bool row;
    size_t t;   // CPML flag tile (64 x 4 x 1 cells from column 1) of the cell, when it is one
};

template <typename T, bool LIN>
__device__ __forceinline__ Em3Lane em3_lane(const Em3Args<T> &A)
{
    Em3Lane L;
    if constexpr (LIN) {
        const int e = (int)blockIdx.x * 256 + (int)(threadIdx.x + 64 * threadIdx.y);
        L.j = e / A.pitch + 1;
        L.i = e - (L.j - 1) * A.pitch - (int)A.xoff + 1;
        L.k = (int)blockIdx.y + 1;
        L.row = L.j <= A.ny;
        L.t = (size_t)((L.i - 1) >> 6) + (size_t)A.tiles_x * ((size_t)((L.j - 1) >> 2) + (size_t)A.tiles_y * (L.k - 1));
    } else {
        const Em3Block B = em3_block(A);
        L.i = B.bx * 64 + threadIdx.x + 1;
        L.j = B.by * 4 + threadIdx.y + 1;
        L.k = B.bz + 1;
        L.row = B.valid && L.i <= A.pitch - (int)A.xoff && L.j <= A.ny;
        L.t = B.bx + (size_t)A.tiles_x * (B.by + (size_t)A.tiles_y * B.bz);
    }
    return L;
}

// Hx, Hy, Hz of cell (i, j, k), cpmlfdtd.f08:3700-3778, updated in place;
// columns past nx (the row padding) take 0 like the cells outside every nest.
template <typename T>
__device__ __forceinline__ void em3_h_cell(const Em3Args<T> &A, int i, int j, int k, bool pml, bool mst)
{
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int64_t sy = A.pitch, sz = (int64_t)A.pitch * ny;
    const int64_t c = A.xoff + (int64_t)(i - 1) + sy * (j - 1) + sz * (k - 1);
    const bool face = i == 1 || j == 1 || k == 1;
    const T *Ex = A.E[0], *Ey = A.E[1], *Ez = A.E[2];
    if (i >= nx || j == ny || k == nz) {   // outside every H nest
        A.H[0][c] = T(0); A.H[1][c] = T(0); A.H[2][c] = T(0);
        if (mst) {
            A.m_dEy_dz[c] = T(0); A.m_dEz_dy[c] = T(0); A.m_dEx_dz[c] = T(0); A.m_dEz_dx[c] = T(0);
            if (k >= 2) { A.m_dEx_dy[c] = T(0); A.m_dEy_dx[c] = T(0); }
        }
        return;
    }
    const T ex0 = Ex[c], exy = Ex[c + sy], exz = Ex[c + sz];
    const T ey0 = Ey[c], eyx = Ey[c + 1], eyz = Ey[c + sz];
    const T ez0 = Ez[c], ezx = Ez[c + 1], ezy = Ez[c + sy];
    const T h0 = face ? T(0) : A.H[0][c], h1 = face ? T(0) : A.H[1][c], h2 = face ? T(0) : A.H[2][c];
    T hx, hy, hz;
    em3_h_math(A, c, k, pml, true, mst, ex0, ey0, ez0, exy, ezy, eyx, ezx, exz, eyz, h0, h1, h2, hx, hy, hz);
}

template <typename T, bool LIN>
__global__ void __launch_bounds__(256) em3_h_kernel(Em3Args<T> A)
{
    const Em3Lane L = em3_lane<T, LIN>(A);
    const int i = L.i, j = L.j, k = L.k;
    const bool row = L.row;
    // one code path with a per-lane CPML flag: a wavefront holding both kinds
    // of cell (the x faces) issues its field loads once, the CPML lanes add
    // theirs
    const bool pml = row && i >= 1 && i < A.nx && j < A.ny && k < A.nz &&
                     (!in_box(A.box, i, j, k) || (A.any_flags && A.flags[L.t]));
    const bool mst = line_any<T>(pml);
    if (!row) return;
    em3_h_cell(A, i, j, k, pml, mst);
}

// Ex, Ey, Ez of cell (i, j, k), cpmlfdtd.f08:3795-3828, zero faces (and row
// padding), in place; returns the cell's squared |E| for the norm.
template <typename T, bool TAB>
__device__ __forceinline__ double em3_e_cell(const Em3Args<T> &A, int i, int j, int k, bool pml, bool mst)
{
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int64_t sy = A.pitch, sz = (int64_t)A.pitch * ny;
    const int64_t c = A.xoff + (int64_t)(i - 1) + sy * (j - 1) + sz * (k - 1);
    const EPre<T> P = em3_e_pre<T, TAB>(A, c, pml);
    if (i == 1 || i >= nx || j == 1 || j == ny || k == 1 || k == nz)   // no H neighbour loads on faces
        return em3_e_new<T, TAB, true>(A, i, j, k, P, T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), mst);
    const T *Hx = A.H[0], *Hy = A.H[1], *Hz = A.H[2];
    const T hx0 = Hx[c], hxy = Hx[c - sy], hxz = Hx[c - sz];
    const T hy0 = Hy[c], hyx = Hy[c - 1], hyz = Hy[c - sz];
    const T hz0 = Hz[c], hzx = Hz[c - 1], hzy = Hz[c - sy];
    return em3_e_new<T, TAB, true>(A, i, j, k, P, hx0, hy0, hz0, hxy, hyx, hzx, hzy, hxz, hyz, mst);
}

template <typename T, bool TAB, bool LIN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) em3_e_kernel(Em3Args<T> A)
{
    const Em3Lane L = em3_lane<T, LIN>(A);
    const int i = L.i, j = L.j, k = L.k;
    const bool row = L.row;
    // CPML cells of the E nests (the faces take 0 and no memory variables)
    const bool pml = row && i > 1 && i < A.nx && j > 1 && j < A.ny && k > 1 && k < A.nz &&
                     (!in_box(A.box, i, j, k) || (A.any_flags && A.flags[L.t]));
    const bool mst = line_any<T>(pml);
    double nrm = 0.0;   // largest squared |E| (sqrt is monotone and correctly rounded)
    if (row) nrm = em3_e_cell<T, TAB>(A, i, j, k, pml, mst);
    block_max_to(&A.vnorm[*A.d_it - 1], sqrt(nrm));
}

// ---- fused step ---------------------------------------------------------------
// One kernel per step: a 64 x TY tile marches kc planes; per plane each thread
// updates H of its cell (cpmlfdtd.f08:3700-3778) into LDS and memory, the
// tile also recomputes H of the column left of it and of the row below it
// (the E update reads H at i-1 and j-1), and after one barrier each thread
// updates E of its cell (3795-3828) from H of this plane (LDS) and of the
// plane below (registers).  E, H and the H-update memory variables are read
// from the previous step's buffers and written to the other ones, so a
// halo recomputation never sees a value a neighbouring tile already
// advanced.  Per cell and step the fields are read once and written once
// (the two half-step kernels read H and E twice).
// TY rows of 64 cells per tile, TY wavefronts: wavefront 0 also updates the
// row below the tile, wavefront 1 (its first TY lanes) the column left of it.
template <typename T, bool TAB, int TY>
__global__ void __launch_bounds__(64 * TY) em3_fused_kernel(Em3Args<T> A, int kc)
{
    __shared__ T hs[2][3][TY + 1][65];   // [plane parity][component][row j0-1 .. j0+TY-1][column i0-1 .. i0+63]
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = blockIdx.x * 64 + 1, j0 = blockIdx.y * TY + 1;
    const int i = i0 + tx, j = j0 + ty;
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int64_t sy = A.pitch, sz = (int64_t)A.pitch * ny;
    const int kb = blockIdx.z * kc + 1, ke = min(kb + kc - 1, nz);
    const bool valid = i <= nx && j <= ny;
    // the halo cell this thread updates besides its own: row j0-1 (wavefront
    // 0) or column i0-1 (lanes 0..TY-1 of wavefront 1)
    const bool hrow = ty == 0 && i <= nx && j0 > 1;
    const bool hcol = ty == 1 && tx < TY && j0 + tx <= ny && i0 > 1;
    const int hi = hrow ? i : i0 - 1, hj = hrow ? j0 - 1 : j0 + tx;
    T hxm = T(0), hym = T(0), unused;
    if (valid && kb > 1) em3_h_new(A, i, j, kb - 1, false, hxm, hym, unused);
    double nrm = 0.0;   // largest squared |E| (sqrt is monotone and correctly rounded)
    for (int k = kb; k <= ke; ++k) {
        const int b = (k - kb) & 1;
        const int64_t c = A.xoff + (int64_t)(i - 1) + sy * (j - 1) + sz * (k - 1);
        const size_t t = (size_t)((i - 1) >> 6) + (size_t)A.tiles_x * ((size_t)((j - 1) >> 2) + (size_t)A.tiles_y * (k - 1));
        const bool pml = !in_box(A.box, i, j, k) || (A.any_flags && A.flags[t]);
        T hx = T(0), hy = T(0), hz = T(0);
        if (valid) em3_h_new(A, i, j, k, true, hx, hy, hz);
        EPre<T> P;
        if (valid) P = em3_e_pre<T, TAB>(A, c, pml);
        hs[b][0][ty + 1][tx + 1] = hx; hs[b][1][ty + 1][tx + 1] = hy; hs[b][2][ty + 1][tx + 1] = hz;
        if (hrow || hcol) {
            T a, cc, e;
            em3_h_new(A, hi, hj, k, false, a, cc, e);
            if (hrow) { hs[b][0][0][tx + 1] = a; hs[b][1][0][tx + 1] = cc; hs[b][2][0][tx + 1] = e; }
            else { hs[b][0][tx + 1][0] = a; hs[b][1][tx + 1][0] = cc; hs[b][2][tx + 1][0] = e; }
        }
        if (ty == 0 && !hrow) { hs[b][0][0][tx + 1] = T(0); hs[b][1][0][tx + 1] = T(0); hs[b][2][0][tx + 1] = T(0); }
        __syncthreads();
        if (valid) {
            const T hxy = hs[b][0][ty][tx + 1], hzy = hs[b][2][ty][tx + 1];
            const T hyx = hs[b][1][ty + 1][tx], hzx = hs[b][2][ty + 1][tx];
            nrm = fmax_nan3<T>(nrm, em3_e_new<T, TAB>(A, i, j, k, P, hx, hy, hz, hxy, hyx, hzx, hzy, hxm, hym));
        }
        hxm = hx;
        hym = hy;
    }
    block_max_to(&A.vnorm[*A.d_it - 1], sqrt(nrm));
}

template <typename T>
static Em3Args<T> em3_args(Device &d)
{
    Em3Args<T> A;
    for (int q = 0; q < 3; ++q) {
        A.E[q] = (T *)d.field[SDT_F_EX + q];
        A.H[q] = (T *)d.field[SDT_F_HX + q];
    }
    const T **mp[] = {&A.aEx, &A.bEx, &A.cEx, &A.bEy, &A.cEy, &A.cEz,
                      &A.s11, &A.s12, &A.s13, &A.s22, &A.s23, &A.s33};
    for (int q = 0; q < 12; ++q) *mp[q] = (const T *)d.mat[q];
    A.mid = d.mid;
    A.tab = (const T *)d.tab;
    A.a = (const T *)d.cpml[0];
    A.b = (const T *)d.cpml[1];
    A.ik = (const T *)d.cpml[2];
    A.ah = (const T *)d.cpml[3];
    A.bh = (const T *)d.cpml[4];
    A.ikh = (const T *)d.cpml[5];
    A.cid = d.cpml_cid;
    A.ctab = (const T *)d.cpml_ctab;
    T **mm[] = {&A.m_dEy_dz, &A.m_dEz_dy, &A.m_dEx_dz, &A.m_dEz_dx, &A.m_dEx_dy, &A.m_dEy_dx,
                &A.m_dHz_dy, &A.m_dHy_dz, &A.m_dHz_dx, &A.m_dHx_dz, &A.m_dHx_dy, &A.m_dHy_dx};
    for (int q = 0; q < 12; ++q) *mm[q] = (T *)d.mem[q];
    A.flags = d.tile_flags;
    A.any_flags = d.nflagged > 0 ? 1 : 0;
    A.col_order = 0;
    A.lin = 0;
    for (int q = 0; q < 6; ++q) A.box[q] = d.box[q];
    A.nx = d.nx; A.ny = d.ny; A.nz = d.nz;
    A.pitch = d.pitch; A.xoff = d.xoff;
    A.tiles_x = d.tiles_x; A.tiles_y = d.tiles_y;
    A.rdx = (T)(1.0 / d.dx);
    A.rdy = (T)(1.0 / d.dy);
    A.rdz = (T)(1.0 / d.dz);
    A.dt = (T)d.dt;
    A.dbH = (T)(d.dt / kMu0);   // cpmlfdtd.f08:3510-3518
    A.d_it = d.d_it;
    A.vnorm = d.norms;
    A.src_cell = d.tail.vfix[0];
    A.pwf = d.pw.faces();
    return A;
}

template <typename T>
static void em3_fused_step_t(Device &d, bool timed)
{
    Em3Args<T> A = em3_args<T>(d);
    for (int q = 0; q < 3; ++q) {
        A.Eo[q] = (const T *)d.field[SDT_F_EX + q];
        A.Ho[q] = (const T *)d.field[SDT_F_HX + q];
        A.E[q] = (T *)d.alt[SDT_F_EX + q];
        A.H[q] = (T *)d.alt[SDT_F_HX + q];
    }
    T **mn[] = {&A.m_dEy_dz, &A.m_dEz_dy, &A.m_dEx_dz, &A.m_dEz_dx, &A.m_dEx_dy, &A.m_dEy_dx};
    for (int q = 0; q < 6; ++q) {
        A.mo[q] = (const T *)d.mem[q];
        *mn[q] = (T *)d.mem[kAltMem + q];
    }
    const int kc = d.zchunk > 0 ? d.zchunk : kEm3FusedChunk;
    const int ty = d.tile_rows == 4 || d.tile_rows == 16 ? d.tile_rows : 8;
    const dim3 grid(d.tiles_x, (d.ny + ty - 1) / ty, (d.nz + kc - 1) / kc), block(64, ty, 1);
    if (timed) (void)hipEventRecord(d.ev[0], d.stream);
    if (ty == 4) {
        if (d.mat_table) hipLaunchKernelGGL((em3_fused_kernel<T, true, 4>), grid, block, 0, d.stream, A, kc);
        else hipLaunchKernelGGL((em3_fused_kernel<T, false, 4>), grid, block, 0, d.stream, A, kc);
    } else if (ty == 16) {
        if (d.mat_table) hipLaunchKernelGGL((em3_fused_kernel<T, true, 16>), grid, block, 0, d.stream, A, kc);
        else hipLaunchKernelGGL((em3_fused_kernel<T, false, 16>), grid, block, 0, d.stream, A, kc);
    } else {
        if (d.mat_table) hipLaunchKernelGGL((em3_fused_kernel<T, true, 8>), grid, block, 0, d.stream, A, kc);
        else hipLaunchKernelGGL((em3_fused_kernel<T, false, 8>), grid, block, 0, d.stream, A, kc);
    }
    if (timed) {
        (void)hipEventRecord(d.ev[1], d.stream);
        (void)hipEventRecord(d.ev[2], d.stream);
    }
    // the new buffers become the current ones
    for (int f = SDT_F_EX; f <= SDT_F_HZ; ++f) {
        void *t = d.field[f];
        d.field[f] = d.alt[f];
        d.alt[f] = t;
    }
    for (int q = 0; q < 6; ++q) {
        void *t = d.mem[q];
        d.mem[q] = d.mem[kAltMem + q];
        d.mem[kAltMem + q] = t;
    }
    d.parity ^= 1;
    launch_plane_wave(d);
    launch_tail(d);
}

template <typename T>
static void em3_step_t(Device &d, bool timed)
{
    if (d.fused) {
        em3_fused_step_t<T>(d, timed);
        return;
    }
    Em3Args<T> A = em3_args<T>(d);
    for (int q = 0; q < 3; ++q) {   // updates in place
        A.Eo[q] = A.E[q];
        A.Ho[q] = A.H[q];
    }
    const T *mp[] = {A.m_dEy_dz, A.m_dEz_dy, A.m_dEx_dz, A.m_dEz_dx, A.m_dEx_dy, A.m_dEy_dx};
    for (int q = 0; q < 6; ++q) A.mo[q] = mp[q];
    A.col_order = d.xcd_remap & 1;
    // linear workgroups in FP64 (404x404x212: 25.5 -> 26.2 GCell/s); in FP32
    // the E update ran slower with them (49.3 -> 46.7 GCell/s), so the 64 x 4
    // tiles stay
    A.lin = !A.col_order && sizeof(T) == 8;
    dim3 grid(d.tiles_x, d.tiles_y, d.tiles_z);
    if (A.lin) grid = dim3((unsigned)((d.ny * (int64_t)d.pitch + 255) / 256), d.nz, 1);
    if (A.col_order) grid = dim3(8 * ((d.tiles_x * d.tiles_y + 7) / 8) * d.tiles_z, 1, 1);
    const dim3 block(64, 4, 1);
    if (timed) (void)hipEventRecord(d.ev[0], d.stream);
    if (A.lin) hipLaunchKernelGGL((em3_h_kernel<T, true>), grid, block, 0, d.stream, A);
    else hipLaunchKernelGGL((em3_h_kernel<T, false>), grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[1], d.stream);
    if (A.lin) {
        if (d.mat_table) hipLaunchKernelGGL((em3_e_kernel<T, true, true>), grid, block, 0, d.stream, A);
        else hipLaunchKernelGGL((em3_e_kernel<T, false, true>), grid, block, 0, d.stream, A);
    } else {
        if (d.mat_table) hipLaunchKernelGGL((em3_e_kernel<T, true, false>), grid, block, 0, d.stream, A);
        else hipLaunchKernelGGL((em3_e_kernel<T, false, false>), grid, block, 0, d.stream, A);
    }
    if (timed) (void)hipEventRecord(d.ev[2], d.stream);
    launch_plane_wave(d);
    launch_tail(d);
}

void em3_step(Device &d, bool timed)
{
    if (d.precision == SDT_FP64) em3_step_t<double>(d, timed);
    else em3_step_t<float>(d, timed);
}

// Inverse permittivity by adjugate/determinant, cpmlfdtd.f08:3640-3661, in
// FP64 on the host; stored as the six unique entries.
int em3_setup(Device &d, const double *const *in)
{
    d.tile_x = 64;
    d.tile_y = 4;
    d.tile_z = 1;
    int rc = compute_tile_flags(d, in);
    if (rc) return rc;
    // the kernels' 64-wide tiles start at column 1: put it on a line boundary
    // so that a tile row covers whole lines that no other tile touches
    set_layout(d, 1);
    const int64_t n = d.ncell;
    auto val = [&](int id, int64_t c) { return in[id] ? in[id][c] : 0.0; };
    auto inv_eps = [&](int64_t c, double *co) {
        const double e11 = val(SDT_E11, c) * kEps0, e12 = val(SDT_E12, c) * kEps0, e13 = val(SDT_E13, c) * kEps0;
        const double e22 = val(SDT_E22, c) * kEps0, e23 = val(SDT_E23, c) * kEps0, e33 = val(SDT_E33, c) * kEps0;
        const double det = e11 * (e22 * e33 - e23 * e23) - e12 * (e12 * e33 - e23 * e13) +
                           e13 * (e12 * e23 - e22 * e13);
        co[0] = (e22 * e33 - e23 * e23) / det;
        co[1] = -(e12 * e33 - e23 * e13) / det;
        co[2] = (e12 * e23 - e22 * e13) / det;
        co[3] = (e11 * e33 - e13 * e13) / det;
        co[4] = -(e11 * e23 - e12 * e13) / det;
        co[5] = (e11 * e22 - e12 * e12) / det;
    };
    d.nmat = 0;
    // few distinct materials (layered / blocky models): one-byte ids and a
    // table of (inverse permittivity, conductivity) rows
    rc = d.mat_table_opt ? build_material_table(d, kEmRow, [&](int64_t c, double *r) {
        inv_eps(c, r);
        for (int q = 0; q < 6; ++q) r[6 + q] = val(SDT_S11 + q, c);
    }) : SDT_OK;
    if (rc) return rc;
    if (!d.mat_table) {
        std::vector<double> co[6];
        for (auto &v : co) v.resize((size_t)n);
        double r[6];
        for (int64_t c = 0; c < n; ++c) {
            inv_eps(c, r);
            for (int q = 0; q < 6; ++q) co[q][(size_t)c] = r[q];
        }
        for (int q = 0; q < 6; ++q) {
            rc = upload_grid(d, co[q].data(), &d.mat[d.nmat++]);
            if (rc) return rc;
        }
        std::vector<double> zero;
        for (int id = SDT_S11; id <= SDT_S33; ++id) {
            if (!in[id] && zero.empty()) zero.assign((size_t)n, 0.0);
            rc = upload_grid(d, in[id] ? in[id] : zero.data(), &d.mat[d.nmat++]);
            if (rc) return rc;
        }
    }
    d.nmem = 12;
    for (int q = 0; q < d.nmem; ++q) {
        rc = alloc_grid_zeroed(d, &d.mem[q]);
        if (rc) return rc;
    }
    for (int f = SDT_F_EX; f <= SDT_F_HZ; ++f) {
        if (f <= SDT_F_EZ && in[SDT_IC_X + f]) rc = upload_grid(d, in[SDT_IC_X + f], &d.field[f]);
        else rc = alloc_grid_zeroed(d, &d.field[f]);
        if (rc) return rc;
    }
    // fused step (on request): second buffers of E, H and the H-update
    // memory variables
    d.fused = d.fused_opt == 1;
    if (d.fused) {
        for (int f = SDT_F_EX; f <= SDT_F_HZ; ++f) {
            rc = alloc_grid_zeroed(d, &d.alt[f]);
            if (rc) return rc;
        }
        for (int q = 0; q < 6; ++q) {
            rc = alloc_grid_zeroed(d, &d.mem[kAltMem + q]);
            if (rc) return rc;
        }
    }

    // point source, cpmlfdtd.f08:3909-3914
    const int is = d.isrc, js = d.jsrc, ks = d.ksrc, nx = d.nx, ny = d.ny, nz = d.nz;
    const int64_t hc = (int64_t)(is - 1) + (int64_t)nx * ((int64_t)(js - 1) + (int64_t)ny * (ks - 1));
    const int64_t c = cell_index(d, is, js, ks);
    const int face = (is == 1 || is == nx || js == 1 || js == ny || ks == 1 || ks == nz);
    TailArgs &t = d.tail;
    t.nterms = 0;
    t.terms[t.nterms++] = {c, SDT_F_EX, SDT_SRC_X, d.dt, val(SDT_E11, hc) * kEps0, face};
    t.terms[t.nterms++] = {c, SDT_F_EY, SDT_SRC_Y, d.dt, val(SDT_E22, hc) * kEps0, face};
    t.terms[t.nterms++] = {c, SDT_F_EZ, SDT_SRC_Z, d.dt, val(SDT_E33, hc) * kEps0, face};
    t.nvfix = 1;
    t.vfix[0] = c;
    t.nvcomp = 3;
    t.vcomp[0] = SDT_F_EX; t.vcomp[1] = SDT_F_EY; t.vcomp[2] = SDT_F_EZ;
    t.stress_norm = 0;
    t.nsfix = 0;
    t.rec_field[0] = SDT_F_EX; t.rec_field[1] = SDT_F_EY; t.rec_field[2] = SDT_F_EZ;

    return upload_cpml_dict(d, in);
}

}  // namespace sdt
