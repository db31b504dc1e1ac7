// Pieces shared by the red-black SOR sweep kernels (kernels.hip,
// sweep_stream.hip, sweep_reg.hip): global-address accessors, vector types
// and the per-cell update.  Every sweep kernel evaluates the update with
// sor_cell, so tile shapes and kernels change the launch pattern, never a
// result bit.
#pragma once
#include "common.h"

namespace iff {

// Tile of this workgroup: blocks are dealt round-robin over the 8 XCDs, so
// block b runs on XCD b & 7; consecutive tiles (neighbours in the grid) are
// mapped to the same XCD and share its L2.  ntiles_pad is a multiple of 8.
__device__ __forceinline__ int tile_of_block(int ntiles_pad)
{
    const int b = blockIdx.x;
    const int per = ntiles_pad >> 3;
    return (b & 7) * per + (b >> 3);
}

// Global-memory accessors: pointers read from the case descriptor are
// generic; casting them to the global address space makes the compiler
// emit global_load/store instead of flat accesses (which also count against
// the LDS wait counter).
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *G(const T *p)
{
    return (const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *GW(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

// Native (clang ext_vector_type) vectors for 2-, 8- and 16-byte accesses.
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef unsigned char u8x2 __attribute__((ext_vector_type(2)));
typedef unsigned char u8x4 __attribute__((ext_vector_type(4)));
template <typename T, int V> struct VecT;
template <> struct VecT<float, 2> { using type = f32x2; };
template <> struct VecT<float, 4> { using type = f32x4; };
template <> struct VecT<double, 2> { using type = f64x2; };
template <int V> struct VecB;
template <> struct VecB<2> { using type = u8x2; };
template <> struct VecB<4> { using type = u8x4; };

template <typename VT>
__device__ __forceinline__ VT gload_vec(const void *p)
{
    return *(const __attribute__((address_space(1))) VT *)p;
}
template <typename VT>
__device__ __forceinline__ void gstore_vec(void *p, VT v)
{
    *(__attribute__((address_space(1))) VT *)p = v;
}

// SOR update of one cell of the 5-point Poisson operator.  aE..aS are the
// face coefficients (zero on closed faces), sa = (aE + aW) + (aN + aS) =
// -ap, r the right side, om the relaxation factor, i the cell's column.
// FP64: the oracle's expression and order, no contraction, IEEE division:
//   sum = (aE pE + aW pW) + (aN pN + aS pS);  w = om / sa;
//   p' = (1 - om) pc + w (sum - r)
// (= pc + om (gs - pc) with gs = (r - sum) / ap in exact arithmetic); the
// weight w depends only on the cell's flags and the metric, so a kernel may
// form it once per launch (sor_weight64) and apply sor_cell_w64 per sweep
// with the same bits.  The FP64 sweeps are bit-identical to the oracle.  FP32 (the error equation
// of the mixed path, refined in FP64 afterwards): the same update
// rearranged for fused multiply-adds and the hardware reciprocal, with the
// weight w = om / sa and wr = w r:
//   s = aS pS + (aN pN + (a2 p2 + a1 p1));  p' = w s + ((1 - om) pc - wr)
// where (a1, p1) is the west neighbour in an odd column and the east
// neighbour in an even one: the neighbour the register-tile kernel takes
// from the adjacent lane, so that the lane shift folds into the product.
// Every FP32 sweep kernel evaluates exactly this.
// FP32 relaxation weight om / sa of a cell (sa != 0).
__device__ __forceinline__ float sor_weight(float sa, float om) { return om * __builtin_amdgcn_rcpf(sa); }

// FP32 update of a case whose face coefficients are one value per axis
// (aE = aW = ax, aN = aS = ay on every face that can open; CaseDev::uni):
// with K = w ax, wr = K r' (r' = r / ax, the form in which the residual
// is stored for such a case) and ratio = ay / ax (one value per case),
//   s = ratio pN + (ratio pS + (pW + pE));  p' = K s + ((1 - om) pc - wr)
// five operations where the general form takes six (one accumulation
// chain, so no more registers live per row than the general form).  A closed face's
// neighbour is a non-fluid or ghost cell, whose error is zero (each solve
// starts from zero and never updates it), so the plain sums equal the
// flagged ones.  Every FP32 sweep kernel evaluates exactly this for such a
// case, and the general form for every other.
__device__ __forceinline__ float sor_cell_u(float pc, float wr, float pW, float pE, float pN, float pS, float K,
                                            float ratio, float c1)
{
    const float s = __builtin_fmaf(ratio, pN, __builtin_fmaf(ratio, pS, pW + pE));
    return __builtin_fmaf(K, s, __builtin_fmaf(c1, pc, -wr));
}
// The same from the cell's sa and r' (kernels that form the weight per update).
__device__ __forceinline__ float sor_cell_u_sa(float pc, float rs, float pE, float pW, float pN, float pS, float sa,
                                               float om, float ax, float ratio)
{
    const float K = sor_weight(sa, om) * ax;
    return sor_cell_u(pc, K * rs, pW, pE, pN, pS, K, ratio, 1.0f - om);
}

// FP64 relaxation weight and update (IEEE division; c1 = 1 - om).
__device__ __forceinline__ double sor_weight64(double sa, double om) { return om / sa; }
__device__ __forceinline__ double sor_cell_w64(double pc, double r, double sum, double w, double c1)
{
    return c1 * pc + w * (sum - r);
}

// FP32 update with w = sor_weight(sa, om), wr = w * r and c1 = 1 - om formed.
__device__ __forceinline__ float sor_cell_w(float pc, float wr, float a1, float p1, float a2, float p2, float aN,
                                            float pN, float aS, float pS, float w, float c1)
{
    const float s = __builtin_fmaf(aS, pS, __builtin_fmaf(aN, pN, __builtin_fmaf(a2, p2, a1 * p1)));
    return __builtin_fmaf(w, s, __builtin_fmaf(c1, pc, -wr));
}

template <typename T>
__device__ __forceinline__ T sor_cell(T pc, T r, T aE, T pE, T aW, T pW, T aN, T pN, T aS, T pS, T sa,
                                      T om, int i)
{
    if constexpr (sizeof(T) == 8) {
        const T sum = (aE * pE + aW * pW) + (aN * pN + aS * pS);
        return sor_cell_w64(pc, r, sum, sor_weight64(sa, om), 1.0 - om);
    } else {
        const float w = sor_weight(sa, om);
        const bool odd = (i & 1) != 0;
        return sor_cell_w(pc, w * r, odd ? aW : aE, odd ? pW : pE, odd ? aE : aW, odd ? pE : pW, aN, pN, aS, pS,
                          w, 1.0f - om);
    }
}

// DPP wave shifts (GFX9 family): lane l receives lane l-1 (from_west) or
// lane l+1 (from_east); the edge lane receives 0 (bound_ctrl), so no
// old value has to be set up first.
__device__ __forceinline__ int dpp_w(int x) { return __builtin_amdgcn_mov_dpp(x, 0x138, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_e(int x) { return __builtin_amdgcn_mov_dpp(x, 0x130, 0xF, 0xF, true); }

__device__ __forceinline__ float from_west(float x) { return __int_as_float(dpp_w(__float_as_int(x))); }
__device__ __forceinline__ float from_east(float x) { return __int_as_float(dpp_e(__float_as_int(x))); }
__device__ __forceinline__ double from_west(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = dpp_w((int)(b & 0xffffffffLL)), hi = dpp_w((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double from_east(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = dpp_e((int)(b & 0xffffffffLL)), hi = dpp_e((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

}  // namespace iff
