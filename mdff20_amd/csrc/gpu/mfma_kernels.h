// mfma_kernels.h -- arithmetic modes and matrix-instruction helpers of the
// matrix-core structure-factor sums (gemm_kernels.h).
//
// S_c(nx,ny,nz) = sum_j w_cj ex_j[nx] ey_j[ny] ez_j[nz] is a complex GEMM
//   S[(nx,c), (ny,nz)] = sum_j A[(nx,c), j] B[j, (ny,nz)],
//   A = w_c ex[nx] (per ion), B = ey[ny] ez[nz] (per ion),
// embedded in a real GEMM with K = 2 x ions (re, im):
//   row (nx,c,re): A' = ( w exr, -w exi ) ; row (nx,c,im): A' = ( w exi, w exr )
//   B'(j,0) = Re(ey ez), B'(j,1) = Im(ey ez).
// Three arithmetic modes:
//   KM_F64    v_mfma_f64_16x16x4f64, FP64 operands and accumulators
//   KM_F32    v_mfma_f32_16x16x4f32, FP32 operands and accumulators
//   KM_F16X3  v_mfma_f32_16x16x32_f16 on split operands x = hi + lo
//             (hi = f16(x), lo = f16(x - hi)) scaled into [-1, 1];
//             A.B ~ Ah.Bh + Ah.Bl + Al.Bh with FP32 accumulation, about 22
//             significant bits per product.
// Result lane layouts (checked on gfx950): f32/f16 16x16 D[4(l/16)+i][l%16],
// f64 16x16x4 D[(l/16)+4i][l%16]; operands A[l%16][k], B[k][l%16] with
// k = l/16 (x4) or 8(l/16)..+7 (x32).
#pragma once
#include <type_traits>

#include "dev.h"

enum { KM_F64 = 0, KM_F32 = 2, KM_F16X3 = 3 };

typedef double mf_d4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef _Float16 mf_h8 __attribute__((ext_vector_type(8)));

#define MF_MT 9    // max row tiles of the forward sum: (Kx+1) * C * 2 <= 144
#define MB_CT 13   // max column tiles of the back-projection: (Kx+1) * 6 * 2 <= 208

// scalar operand type of the non-split modes
template <int MODE>
using mf_sc_t = typename std::conditional<MODE == KM_F64, double, float>::type;

template <int MODE>
__device__ __forceinline__ auto mf_mma(mf_sc_t<MODE> a, mf_sc_t<MODE> b,
                                       typename std::conditional<MODE == KM_F64, mf_d4, mf_f4>::type c)
{
    if constexpr (MODE == KM_F64) return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// the back-projection's moments {1, ny, nz, ny^2, ny nz, nz^2} of a point
__device__ __forceinline__ double mb_moment(int m, int ny, int nz)
{
    switch (m) {
    case 0: return 1.0;
    case 1: return (double)ny;
    case 2: return (double)nz;
    case 3: return (double)(ny * ny);
    case 4: return (double)(ny * nz);
    default: return (double)(nz * nz);
    }
}
