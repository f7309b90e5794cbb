// clm_device.hpp -- device data layout and shared device helpers (gfx950, wave64).
//
// Layout in HBM after a cell-list update (one context):
//   E[ncap]      entries sorted by computing cell, 16 B (f32) / 32 B (f64) each:
//                x, y, z in the aligned frame + w = index | type<<28 | ghost<<31.
//                Within a cell the real particles come first, then ghost images
//                (the reference keeps both in one per-cell vector, CellLists.jl:67-82).
//   cstart[ncells+1]  first entry of each cell (exclusive scan of the counts).
//   creal[ncells]     number of real particles of each cell.
//   hstart[ncells+1]  exclusive scan of the per-cell "home" counts of the sweep mode.
//   tstart[nrows+1]   exclusive scan of the per-row wave-task counts.
// Cells are numbered column-major (x fastest, CellOperations.jl:256-257), so each
// x-row of cells (fixed y, z) is one contiguous range of E.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clm {

// Entry word: particle index << 4 | type << 1 | ghost.  With the index in the high bits,
// comparing whole words orders entries of different particles by index.
constexpr uint32_t IDX_LIMIT = 1u << 28;
__host__ __device__ inline uint32_t w_make(uint32_t idx, uint32_t type, bool ghost) {
    return (idx << 4) | ((type & 7u) << 1) | (ghost ? 1u : 0u);
}
__host__ __device__ inline uint32_t w_idx(uint32_t w) { return w >> 4; }
__host__ __device__ inline uint32_t w_type(uint32_t w) { return (w >> 1) & 7u; }
__host__ __device__ inline bool w_ghost(uint32_t w) { return (w & 1u) != 0; }

template <typename T> struct Ent;
template <> struct alignas(16) Ent<float> { float x, y, z; uint32_t w; };
template <> struct alignas(32) Ent<double> { double x, y, z; uint64_t w; };

template <typename T>
__host__ __device__ inline uint32_t ent_w(const Ent<T> &e) { return (uint32_t)e.w; }

// Box parameters the device kernels need (all values in the working precision).
template <typename T>
struct DevBox {
    T input[9];       // column-major, input unit cell (wrap)
    T rot[9];         // rotation applied after wrapping (identity unless triclinic)
    T shift[27][3];   // aligned * k for the 27 k in {-1,0,1}^3, Iterators.product order
    T cbmin[3], cbmax[3], cs[3];
    T rc2;
    int nc[3];
    int lcell;
    int celltype;
    int dim;
};

// Flags raised by the binning kernels (device word, read back by the host).
enum : uint32_t { FLAG_NAN = 1u, FLAG_RANGE = 2u, FLAG_OVERFLOW = 4u, FLAG_SCAN = 8u };

__device__ inline int lane_id() { return (int)__lane_id(); }

template <typename T>
__device__ inline T dev_floor(T v);
template <>
__device__ inline float dev_floor<float>(float v) { return floorf(v); }
template <>
__device__ inline double dev_floor<double>(double v) { return floor(v); }

// Correctly rounded square root and division.  The translation unit is compiled with
// -fhip-fp32-correctly-rounded-divide-sqrt, under which llvm.sqrt.f32 and fdiv are
// lowered to IEEE round-to-nearest sequences (HIP's __fsqrt_rn maps to the native,
// 1-ulp instruction unless OCML_BASIC_ROUNDED_OPERATIONS is defined, so it is not used).
template <typename T>
__device__ inline T dev_sqrt_rn(T v);
template <>
__device__ inline float dev_sqrt_rn<float>(float v) { return __builtin_sqrtf(v); }
template <>
__device__ inline double dev_sqrt_rn<double>(double v) { return __builtin_sqrt(v); }

// Correctly rounded square root of a float x >= 2^-96: the hardware estimate (within an
// ulp) corrected by one step to either side -- the candidate whose residual x - s*s'
// (one fma) changes sign.  For such x this is the full routine's result; the full
// routine also scales smaller arguments and passes zero, infinities and NaN through.
__device__ inline float sqrt_rn_above_tiny(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const int si = __builtin_bit_cast(int, s);
    const float sm = __builtin_bit_cast(float, si - 1), sp = __builtin_bit_cast(float, si + 1);
    float r = fmaf(-sm, s, x) <= 0.0f ? sm : s;
    r = fmaf(-sp, s, x) > 0.0f ? sp : r;
    return r;
}
constexpr float SQRT_FAST_MIN = 0x1p-96f;

template <typename T>
__device__ inline T dev_div_rn(T a, T b) { return a / b; }

template <typename T>
__device__ inline bool dev_finite(T v) { return isfinite(v); }

// Reciprocal for tolerance-level arithmetic (energies, forces): the hardware estimate
// refined by Newton steps -- one for float (~1 ulp), two for double (~1 ulp).
template <typename T>
__device__ inline T fast_rcp(T x);
template <>
__device__ inline float fast_rcp<float>(float x) {
    float r = __builtin_amdgcn_rcpf(x);
    return fmaf(fmaf(-x, r, 1.0f), r, r);
}
template <>
__device__ inline double fast_rcp<double>(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}

// Wrapping to the first cell followed by the rotation (CellOperations.jl:56-94,
// CellLists.jl:944-945), 3x3 Cramer solve in StaticArrays order.  The translation
// unit is compiled with -ffp-contract=off so every product and sum rounds on its own.
template <typename T>
__device__ inline void wrap_rotate3(const DevBox<T> &b, const T *x, T *out) {
    const T *a = b.input;
#define A(i, j) a[(i) + 3 * (j)]
    T cc0 = A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2);
    T cc1 = A(2, 1) * A(0, 2) - A(0, 1) * A(2, 2);
    T cc2 = A(0, 1) * A(1, 2) - A(1, 1) * A(0, 2);
    T d = (A(0, 0) * cc0 + A(1, 0) * cc1) + A(2, 0) * cc2;
    T p0 = (((A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) * x[0] + (A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2)) * x[1]) +
            (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) * x[2]) / d;
    T p1 = (((A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2)) * x[0] + (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) * x[1]) +
            (A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2)) * x[2]) / d;
    T p2 = (((A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0)) * x[0] + (A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1)) * x[1]) +
            (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0)) * x[2]) / d;
    T p[3] = {p0, p1, p2};
    for (int k = 0; k < 3; ++k) {
        p[k] = p[k] - dev_floor(p[k]);
        if (p[k] == T(1)) p[k] = T(0);
    }
    T w[3];
    for (int i = 0; i < 3; ++i) w[i] = (A(i, 0) * p[0] + A(i, 1) * p[1]) + A(i, 2) * p[2];
#undef A
    if (b.celltype == 1) {
        const T *r = b.rot;
        for (int i = 0; i < 3; ++i) out[i] = (r[i] * w[0] + r[i + 3] * w[1]) + r[i + 6] * w[2];
    } else {
        for (int i = 0; i < 3; ++i) out[i] = w[i];
    }
}

template <typename T>
__device__ inline void wrap_rotate2(const DevBox<T> &b, const T *x, T *out) {
    const T *a = b.input;
    T a11 = a[0], a21 = a[1], a12 = a[3], a22 = a[4];
    T d = a11 * a22 - a12 * a21;
    T p[2] = {(a22 * x[0] - a12 * x[1]) / d, (a11 * x[1] - a21 * x[0]) / d};
    for (int k = 0; k < 2; ++k) {
        p[k] = p[k] - dev_floor(p[k]);
        if (p[k] == T(1)) p[k] = T(0);
    }
    T w[2] = {a11 * p[0] + a12 * p[1], a21 * p[0] + a22 * p[1]};
    if (b.celltype == 1) {
        const T *r = b.rot;
        out[0] = r[0] * w[0] + r[3] * w[1];
        out[1] = r[1] * w[0] + r[4] * w[1];
    } else {
        out[0] = w[0];
        out[1] = w[1];
    }
}

// particle_cell (Box.jl:499-508), 0-based; real particles get the border rule of
// CellLists.jl:956-967 and are then kept in [lcell, nc-lcell-1] (a particle outside
// that range would make the reference raise a BoundsError; it raises FLAG_RANGE here).
template <typename T>
__device__ inline int cell_of(const DevBox<T> &b, const T *q, bool real, uint32_t &flags) {
    int c[3] = {0, 0, 0};
    const int L = b.lcell;
    for (int k = 0; k < b.dim; ++k) {
        T f = dev_floor((q[k] - b.cbmin[k]) / b.cs[k]);
        int ck;
        if (!(f > T(-2)))
            ck = -1;
        else if (f > T(b.nc[k] + 1))
            ck = b.nc[k] + 1;
        else
            ck = (int)f;  // 0-based == reference index - 1
        if (real) {
            if (ck == L - 1) ck += 1;
            if (ck == b.nc[k] - L) ck -= 1;
            if (ck < L || ck > b.nc[k] - L - 1) {
                flags |= FLAG_RANGE;
                ck = ck < L ? L : b.nc[k] - L - 1;
            }
        } else {
            if (ck < 0 || ck > b.nc[k] - 1) {
                flags |= FLAG_RANGE;
                ck = ck < 0 ? 0 : b.nc[k] - 1;
            }
        }
        c[k] = ck;
    }
    return c[0] + b.nc[0] * (c[1] + b.nc[1] * c[2]);
}

template <typename T>
__device__ inline bool in_box(const DevBox<T> &b, const T *q) {
    bool ok = true;
    for (int k = 0; k < b.dim; ++k) ok = ok && (b.cbmin[k] <= q[k]) && (q[k] < b.cbmax[k]);
    return ok;
}

// 64-lane helpers
__device__ inline uint64_t ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ inline int rank_in(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ inline int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <typename T>
__device__ inline T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace clm
