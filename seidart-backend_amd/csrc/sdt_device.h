// sdt_device.h -- device-side helpers shared by the gfx950 CPML-FDTD kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sdt {

// Fortran-order linear index of 1-based (i, j, k) on an nx*ny*nz grid.
__device__ __forceinline__ int64_t lin3(int i, int j, int k, int nx, int ny)
{
    return (int64_t)(i - 1) + (int64_t)nx * ((int64_t)(j - 1) + (int64_t)ny * (int64_t)(k - 1));
}

// scalar_mean(a, b, code) of averaging.f08:5-25 (1 harmonic, 2 geometric,
// 3 arithmetic, 4 none).
template <typename T>
__device__ __forceinline__ T scalar_mean(T a, T b, int code)
{
    switch (code) {
    case 1: return T(2) * a * b / (a + b);
    case 2: return sqrt(a * b);
    case 3: return T(0.5) * (a + b);
    default: return a;
    }
}

// Voigt slot q (c11 c12 c13 c14 c15 c16 c22 ... c66 order) of an isotropic-form
// tensor: c11 = c22 = c33, c12 = c13 = c23, c44 = c55 = c66, the rest 0.  With
// q a compile-time constant after unrolling, only the three arrays are read.
template <typename T>
__device__ __forceinline__ T iso_entry(int q, T c11, T c12, T c44)
{
    return (q == 0 || q == 6 || q == 11) ? c11
         : (q == 1 || q == 2 || q == 7)   ? c12
         : (q == 15 || q == 18 || q == 20) ? c44
                                           : T(0);
}

// Max of non-negative doubles through their IEEE bit patterns (monotone for
// x >= 0; NaN patterns compare above +inf, so a NaN norm reads as a blow-up).
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *slot, double v)
{
    atomicMax(slot, (unsigned long long)__double_as_longlong(v));
}

// Workgroup max-reduction of a non-negative value followed by at most one
// atomic per workgroup: the slot is read first (agent-scope atomic load, past
// the CU's L1) and the atomic is skipped when it cannot raise the value.  Once
// a wave fills the grid every workgroup has a non-zero maximum, and one atomic
// from each of ~1e5 workgroups on a single address serialises at L2 (it
// doubled the 3D EM E update); a stale read only costs an extra atomic.
// blockDim.x*blockDim.y*blockDim.z must be a multiple of 64.
// Whole-line stores.  A 128-byte line that a kernel writes only in part costs
// the memory a read of the line besides the write.  Where the lanes of a
// wavefront store to consecutive elements, lane 0 on a line boundary, each
// group of 128 / sizeof(T) lanes shares one line: line_any tells a lane
// whether any lane of its group has p set (call it with every lane of the
// wavefront active).
template <typename T>
__device__ __forceinline__ bool line_any(bool p)
{
    constexpr int L = 128 / (int)sizeof(T);
    const unsigned long long m = __ballot(p);
    const int base = (int)__lane_id() & ~(L - 1);
    return ((m >> base) & (L >= 64 ? ~0ull : ((1ull << L) - 1))) != 0;
}

__device__ __forceinline__ void block_max_to(unsigned long long *slot, double v)
{
    __shared__ double wave_max[16];
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int lane = tid & 63, wave = tid >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_down(v, off, 64);
        v = (o > v || o != o) ? o : v;
    }
    if (lane == 0) wave_max[wave] = v;
    __syncthreads();
    if (tid == 0) {
        const int nw = (blockDim.x * blockDim.y * blockDim.z + 63) >> 6;
        double m = wave_max[0];
        for (int w = 1; w < nw; ++w) m = (wave_max[w] > m || wave_max[w] != wave_max[w]) ? wave_max[w] : m;
        if (m > 0.0 || m != m) {
            const unsigned long long bits = (unsigned long long)__double_as_longlong(m);
            if (bits > __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(slot, bits);
        }
    }
}

// Streaming loads/stores for data touched once per step (materials, memory
// variables, CPML coefficients): marked non-temporal so they do not evict
// the stencil operands that neighbouring cells re-read from L2.
template <typename T>
__device__ __forceinline__ T ldnt(const T *p)
{
    return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ __forceinline__ void stnt(T *p, T v)
{
    __builtin_nontemporal_store(v, p);
}

__device__ __forceinline__ bool in_box(const int (&b)[6], int i, int j, int k)
{
    return i >= b[0] && i <= b[1] && j >= b[2] && j <= b[3] && k >= b[4] && k <= b[5];
}

// gfx950 (MI355X) has 8 XCDs, each with its own L2; the dispatcher hands
// workgroup b to XCD b % 8.  xcd_remap turns the hardware id into a logical
// id such that every XCD owns one contiguous range of logical ids, so tiles
// that share stencil halos run on the same XCD and hit the same L2.
constexpr int kNumXcd = 8;
__device__ __forceinline__ unsigned xcd_remap(unsigned bid, unsigned nb)
{
    const unsigned x = bid % kNumXcd, s = bid / kNumXcd;
    const unsigned per = nb / kNumXcd, rem = nb % kNumXcd;
    return x * per + (x < rem ? x : rem) + s;
}

// Per-launch constants shared by every solver kernel.
struct GridInfo {
    int nx, ny, nz;
    int tiles_x, tiles_y;   // tile counts of the launch grid (flags indexing)
};

}  // namespace sdt
