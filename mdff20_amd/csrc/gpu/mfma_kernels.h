// mfma_kernels.h -- matrix-core (MFMA) form of the structure-factor sum.
//
// S_c(nx,ny,nz) = sum_j w_cj ex_j[nx] ey_j[ny] ez_j[nz] is a complex GEMM
//   S[(nx,c), (ny,nz)] = sum_j A[(nx,c), j] B[j, (ny,nz)],
//   A = w_c ex[nx] (per ion), B = ey[ny] ez[nz] (per ion),
// embedded in a real GEMM with K = 2 x ions (re, im):
//   row (nx,c,re): A' = ( w exr, -w exi ) ; row (nx,c,im): A' = ( w exi, w exr )
//   B'(j,0) = Re(ey ez), B'(j,1) = Im(ey ez).
// Both operands are generated per block into LDS from the e^{i n theta}
// tables; the ey*ez product is shared by every row.  Three arithmetic modes:
//   KM_F64    v_mfma_f64_16x16x4f64, FP64 operands and accumulators
//   KM_F32    v_mfma_f32_16x16x4f32, FP32 operands and accumulators
//   KM_BF16X3 v_mfma_f32_16x16x32_bf16 on split operands x = hi + lo
//             (hi = bf16(x), lo = bf16(x - hi)); A.B ~ Ah.Bh + Ah.Bl + Al.Bh,
//             FP32 accumulation, about 16 significant bits per product.
//   KM_F16X3  v_mfma_f32_16x16x32_f16 on split operands (hi = f16(x),
//             lo = f16(x - hi)) scaled into [-1, 1], same three products:
//             about 22 bits (stored-operand GEMMs of gemm_kernels.h only).
// Result lane layouts (checked on gfx950): f32/bf16 16x16 D[4(l/16)+i][l%16],
// f64 16x16x4 D[(l/16)+4i][l%16]; operands A[l%16][k], B[k][l%16] with
// k = l/16 (x4) or 8(l/16)..+7 (x32).
#pragma once
#include <type_traits>

#include "dev.h"

enum { KM_F64 = 0, KM_BF16X3 = 1, KM_F32 = 2, KM_F16X3 = 3 };

typedef double mf_d4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 mf_bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 mf_h8 __attribute__((ext_vector_type(8)));

#define MF_MT 9       // max row tiles: (Kx+1) * C * 2 <= 144
#define MF_LDA_PAD 1  // FP64 LDS row padding (in elements)

struct SkMArgs {
    int nsel, Kx, Ky, Kz, P, C, M;
    const void *tx, *ty, *tz;  // tables: double (KM_F64) or float (other modes)
    const int *idx;
    const double *q, *mu;
    int n;
    int chunk;                 // ions per block (multiple of the sub-batch)
    double *part;              // [nchunk][(Kx+1)*P][C][2]
    const int *active;
};

template <int MODE>
struct MfTraits;
template <>
struct MfTraits<KM_F64> {
    using tab_t = double;
    static constexpr int SA = 16;   // ions per LDS sub-batch
    static constexpr int CW = 1;    // column tiles per wave
};
template <>
struct MfTraits<KM_BF16X3> {
    using tab_t = float;
    static constexpr int SA = 32;
    static constexpr int CW = 2;
};
template <>
struct MfTraits<KM_F32> {
    using tab_t = float;
    static constexpr int SA = 16;
    static constexpr int CW = 1;
};

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

__device__ __forceinline__ double mf_weight(const SkMArgs &a, int c, int ion)
{
    if (a.C == 1) return a.q[ion];
    if (a.C == 3) return a.mu[(size_t)c * a.n + ion];
    return c == 0 ? a.q[ion] : a.mu[(size_t)(c - 1) * a.n + ion];
}

template <int MODE>
__global__ void __launch_bounds__(256) k_sk_fwd_mfma(SkMArgs a)
{
    if (a.active && !*a.active) return;
    using T = typename MfTraits<MODE>::tab_t;
    constexpr int SA = MfTraits<MODE>::SA, KS = 2 * SA, CW = MfTraits<MODE>::CW;
    constexpr int BC = 64 * CW;            // columns per block
    constexpr int ROWS = MF_MT * 16;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int NZ = 2 * a.Kz + 1;
    const int col0 = blockIdx.x * BC;
    const int j0 = blockIdx.y * a.chunk;
    const int j1 = min(j0 + a.chunk, a.nsel);
    const int MT = (a.M + 15) / 16;
    const T *tx = (const T *)a.tx, *ty = (const T *)a.ty, *tz = (const T *)a.tz;

    // LDS operand tiles
    constexpr int LDA = (MODE != KM_BF16X3) ? KS + MF_LDA_PAD : KS + 8;
    using sc = mf_sc_t<MODE>;
    sc *As = (sc *)smem_raw;
    sc *Bs = As + ROWS * LDA;
    __bf16 *Ash = (__bf16 *)smem_raw;
    __bf16 *Asl = Ash + ROWS * LDA;
    __bf16 *Bsh = Asl + ROWS * LDA;
    __bf16 *Bsl = Bsh + BC * LDA;

    using acc_t = typename std::conditional<MODE == KM_F64, mf_d4, mf_f4>::type;
    acc_t acc[MF_MT][CW];
#pragma unroll
    for (int t = 0; t < MF_MT; t++)
#pragma unroll
        for (int c = 0; c < CW; c++) acc[t][c] = acc_t{0, 0, 0, 0};

    for (int jb = j0; jb < j1; jb += SA) {
        __syncthreads();
        // B' = Re/Im(ey ez) for the block's columns
        for (int e = threadIdx.x; e < BC * SA; e += 256) {
            int col = e / SA, jj = e % SA;
            int p = col0 + col, j = jb + jj;
            double br = 0.0, bi = 0.0;
            if (p < a.P && j < j1) {
                int iy = p / NZ, iz = p % NZ;
                size_t oy = 2 * ((size_t)iy * a.nsel + j), oz = 2 * ((size_t)iz * a.nsel + j);
                double yr = ty[oy], yi = ty[oy + 1], zr = tz[oz], zi = tz[oz + 1];
                br = yr * zr - yi * zi;
                bi = yr * zi + yi * zr;
            }
            if constexpr (MODE != KM_BF16X3) {
                Bs[col * LDA + 2 * jj] = (sc)br;
                Bs[col * LDA + 2 * jj + 1] = (sc)bi;
            } else {
                float fr = (float)br, fi = (float)bi;
                __bf16 hr = (__bf16)fr, hi_ = (__bf16)fi;
                Bsh[col * LDA + 2 * jj] = hr;
                Bsh[col * LDA + 2 * jj + 1] = hi_;
                Bsl[col * LDA + 2 * jj] = (__bf16)(fr - (float)hr);
                Bsl[col * LDA + 2 * jj + 1] = (__bf16)(fi - (float)hi_);
            }
        }
        // A' rows (nx, c, re/im)
        for (int e = threadIdx.x; e < ROWS * SA; e += 256) {
            int r = e / SA, jj = e % SA;
            int j = jb + jj;
            double v0 = 0.0, v1 = 0.0;
            if (r < a.M && j < j1) {
                int nx = r / (2 * a.C), c = (r >> 1) % a.C, ri = r & 1;
                int ion = a.idx ? a.idx[j] : j;
                double w = mf_weight(a, c, ion);
                size_t ox = 2 * ((size_t)nx * a.nsel + j);
                double xr = tx[ox], xi = tx[ox + 1];
                if (ri == 0) { v0 = w * xr; v1 = -w * xi; }
                else { v0 = w * xi; v1 = w * xr; }
            }
            if constexpr (MODE != KM_BF16X3) {
                As[r * LDA + 2 * jj] = (sc)v0;
                As[r * LDA + 2 * jj + 1] = (sc)v1;
            } else {
                float f0 = (float)v0, f1 = (float)v1;
                __bf16 h0 = (__bf16)f0, h1 = (__bf16)f1;
                Ash[r * LDA + 2 * jj] = h0;
                Ash[r * LDA + 2 * jj + 1] = h1;
                Asl[r * LDA + 2 * jj] = (__bf16)(f0 - (float)h0);
                Asl[r * LDA + 2 * jj + 1] = (__bf16)(f1 - (float)h1);
            }
        }
        __syncthreads();
        if constexpr (MODE != KM_BF16X3) {
            const int k = lane >> 4, cl = lane & 15;
#pragma unroll
            for (int kk = 0; kk < KS / 4; kk++) {
                sc b = Bs[(wave * 16 + cl) * LDA + kk * 4 + k];
#pragma unroll
                for (int t = 0; t < MF_MT; t++) {
                    if (t < MT) {
                        sc av = As[(t * 16 + cl) * LDA + kk * 4 + k];
                        acc[t][0] = mf_mma<MODE>(av, b, acc[t][0]);
                    }
                }
            }
        } else {
            const int kg = (lane >> 4) * 8, cl = lane & 15;
#pragma unroll
            for (int kk = 0; kk < KS / 32; kk++) {
                mf_bf8 bh[CW], bl[CW];
#pragma unroll
                for (int ct = 0; ct < CW; ct++) {
                    int col = (wave * CW + ct) * 16 + cl;
                    bh[ct] = *(const mf_bf8 *)(Bsh + col * LDA + kk * 32 + kg);
                    bl[ct] = *(const mf_bf8 *)(Bsl + col * LDA + kk * 32 + kg);
                }
#pragma unroll
                for (int t = 0; t < MF_MT; t++) {
                    if (t < MT) {
                        mf_bf8 ah = *(const mf_bf8 *)(Ash + (t * 16 + cl) * LDA + kk * 32 + kg);
                        mf_bf8 al = *(const mf_bf8 *)(Asl + (t * 16 + cl) * LDA + kk * 32 + kg);
#pragma unroll
                        for (int ct = 0; ct < CW; ct++) {
                            acc[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[ct], acc[t][ct], 0, 0, 0);
                            acc[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[ct], acc[t][ct], 0, 0, 0);
                            acc[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[ct], acc[t][ct], 0, 0, 0);
                        }
                    }
                }
            }
        }
    }
    // partial sums of this ion chunk, in the layout k_sk_reduce reads
    const size_t nkf = (size_t)(a.Kx + 1) * a.P;
    double *out = a.part + (size_t)blockIdx.y * nkf * (2 * a.C);
#pragma unroll
    for (int t = 0; t < MF_MT; t++) {
        if (t >= MT) continue;
#pragma unroll
        for (int ct = 0; ct < CW; ct++) {
            int p = col0 + (wave * CW + ct) * 16 + (lane & 15);
            if (p >= a.P) continue;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                int r = (MODE == KM_F64) ? t * 16 + (lane >> 4) + 4 * i : t * 16 + 4 * (lane >> 4) + i;
                if (r >= a.M) continue;
                int nx = r / (2 * a.C), c = (r >> 1) % a.C, ri = r & 1;
                out[((size_t)nx * a.P + p) * (2 * a.C) + 2 * c + ri] = (double)acc[t][ct][i];
            }
        }
    }
}

template <int MODE>
inline size_t mf_fwd_lds_bytes()
{
    constexpr int SA = MfTraits<MODE>::SA, KS = 2 * SA, CW = MfTraits<MODE>::CW, BC = 64 * CW;
    constexpr int ROWS = MF_MT * 16;
    if (MODE != KM_BF16X3) return (size_t)(ROWS + BC) * (KS + MF_LDA_PAD) * sizeof(mf_sc_t<MODE>);
    return (size_t)2 * (ROWS + BC) * (KS + 8) * sizeof(__bf16);
}

// ---------------------------------------------------------------------------
// Back-projection on the matrix cores.  With k linear in n, the field and
// field-gradient sums need, per ion j and per nx,
//   T_m(j,nx) = sum_(ny,nz) ey_j[ny] ez_j[nz] W(nx,ny,nz) mom_m(ny,nz),
//   mom = {1, ny, nz, ny^2, ny nz, nz^2}  (first 3 for the field only),
// a complex GEMM with rows = ions, K = (ny,nz) and columns = (nx, m); the
// epilogue forms Z_m = ex_j[nx] T_m and
//   H_x += nx Im Z_0, H_y += Im Z_1, H_z += Im Z_2,
//   M_xx += nx^2 Re Z_0, M_xy += nx Re Z_1, M_xz += nx Re Z_2,
//   M_yy += Re Z_3, M_yz += Re Z_4, M_zz += Re Z_5.
// Real embedding: A'(j,(p,0)) = Re(ey ez), A'(j,(p,1)) = Im(ey ez);
// B'((p,0),(nx,m,re)) = Re Wm, B'((p,1),(nx,m,re)) = -Im Wm,
// B'((p,0),(nx,m,im)) = Im Wm, B'((p,1),(nx,m,im)) = Re Wm.
struct BkMArgs {
    int nsel, Kx, Ky, Kz, P, CH, NC;   // CH moments (3 or 6), NC = (Kx+1)*CH*2
    const void *tx, *ty, *tz;
    const double *W;                   // [(Kx+1)*P][2]
    int psplit;                        // (ny,nz) points per grid.y split
    double *part;                      // [nsplit][3 or 9][nsel]
    const int *active;
};

#define MB_CT 13   // max column tiles: (Kx+1) * 6 * 2 <= 208

template <int MODE>
struct MbTraits;
template <>
struct MbTraits<KM_F64> {
    static constexpr int SP = 16;   // (ny,nz) points per LDS sub-batch (K = 32)
};
template <>
struct MbTraits<KM_BF16X3> {
    static constexpr int SP = 32;   // K = 64
};
template <>
struct MbTraits<KM_F32> {
    static constexpr int SP = 16;
};

template <int MODE>
inline size_t mf_back_lds_bytes()
{
    constexpr int KS = 2 * MbTraits<MODE>::SP;
    constexpr int ROWS = 64, COLS = MB_CT * 16;
    if (MODE != KM_BF16X3) return (size_t)(ROWS + COLS) * (KS + MF_LDA_PAD) * sizeof(mf_sc_t<MODE>);
    return (size_t)2 * (ROWS + COLS) * (KS + 8) * sizeof(__bf16);
}

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

template <int MODE>
__global__ void __launch_bounds__(256) k_sk_back_mfma(BkMArgs a)
{
    if (a.active && !*a.active) return;
    using T = typename MfTraits<MODE>::tab_t;
    constexpr int SP = MbTraits<MODE>::SP, KS = 2 * SP;
    constexpr int ROWS = 64, COLS = MB_CT * 16;
    constexpr int LDA = (MODE != KM_BF16X3) ? KS + MF_LDA_PAD : KS + 8;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int NZ = 2 * a.Kz + 1;
    const int jrow0 = blockIdx.x * ROWS;
    const int p0 = blockIdx.y * a.psplit;
    const int p1 = min(p0 + a.psplit, a.P);
    const int CT = (a.NC + 15) / 16;
    const T *tx = (const T *)a.tx, *ty = (const T *)a.ty, *tz = (const T *)a.tz;
    using sc = mf_sc_t<MODE>;
    sc *As = (sc *)smem_raw;
    sc *Bs = As + ROWS * LDA;
    __bf16 *Ash = (__bf16 *)smem_raw;
    __bf16 *Asl = Ash + ROWS * LDA;
    __bf16 *Bsh = Asl + ROWS * LDA;
    __bf16 *Bsl = Bsh + COLS * LDA;

    using acc_t = typename std::conditional<MODE == KM_F64, mf_d4, mf_f4>::type;
    acc_t acc[MB_CT];
#pragma unroll
    for (int c = 0; c < MB_CT; c++) acc[c] = acc_t{0, 0, 0, 0};

    for (int pb = p0; pb < p1; pb += SP) {
        __syncthreads();
        // A' = Re/Im(ey ez) for the block's 64 ions
        for (int e = threadIdx.x; e < ROWS * SP; e += 256) {
            int r = e / SP, pp = e % SP;
            int j = jrow0 + r, p = pb + pp;
            double br = 0.0, bi = 0.0;
            if (j < a.nsel && p < p1) {
                int iy = p / NZ, iz = p % NZ;
                size_t oy = 2 * ((size_t)iy * a.nsel + j), oz = 2 * ((size_t)iz * a.nsel + j);
                double yr = ty[oy], yi = ty[oy + 1], zr = tz[oz], zi = tz[oz + 1];
                br = yr * zr - yi * zi;
                bi = yr * zi + yi * zr;
            }
            if constexpr (MODE != KM_BF16X3) {
                As[r * LDA + 2 * pp] = (sc)br;
                As[r * LDA + 2 * pp + 1] = (sc)bi;
            } else {
                float fr = (float)br, fi = (float)bi;
                __bf16 hr = (__bf16)fr, hi_ = (__bf16)fi;
                Ash[r * LDA + 2 * pp] = hr;
                Ash[r * LDA + 2 * pp + 1] = hi_;
                Asl[r * LDA + 2 * pp] = (__bf16)(fr - (float)hr);
                Asl[r * LDA + 2 * pp + 1] = (__bf16)(fi - (float)hi_);
            }
        }
        // B' columns (nx, m, re/im), stored per column with K contiguous
        for (int e = threadIdx.x; e < COLS * SP; e += 256) {
            int col = e / SP, pp = e % SP;
            int p = pb + pp;
            double v0 = 0.0, v1 = 0.0;
            if (col < a.NC && p < p1) {
                int nx = col / (2 * a.CH), m = (col >> 1) % a.CH, ri = col & 1;
                int iy = p / NZ, iz = p % NZ;
                double mom = mb_moment(m, iy - a.Ky, iz - a.Kz);
                size_t kf = (size_t)nx * a.P + p;
                double wr = a.W[2 * kf] * mom, wi = a.W[2 * kf + 1] * mom;
                if (ri == 0) { v0 = wr; v1 = -wi; }
                else { v0 = wi; v1 = wr; }
            }
            if constexpr (MODE != KM_BF16X3) {
                Bs[col * LDA + 2 * pp] = (sc)v0;
                Bs[col * LDA + 2 * pp + 1] = (sc)v1;
            } else {
                float f0 = (float)v0, f1 = (float)v1;
                __bf16 h0 = (__bf16)f0, h1 = (__bf16)f1;
                Bsh[col * LDA + 2 * pp] = h0;
                Bsh[col * LDA + 2 * pp + 1] = h1;
                Bsl[col * LDA + 2 * pp] = (__bf16)(f0 - (float)h0);
                Bsl[col * LDA + 2 * pp + 1] = (__bf16)(f1 - (float)h1);
            }
        }
        __syncthreads();
        if constexpr (MODE != KM_BF16X3) {
            const int k = lane >> 4, cl = lane & 15;
#pragma unroll
            for (int kk = 0; kk < KS / 4; kk++) {
                sc av = As[(wave * 16 + cl) * LDA + kk * 4 + k];
#pragma unroll
                for (int ct = 0; ct < MB_CT; ct++) {
                    if (ct < CT) {
                        sc b = Bs[(ct * 16 + cl) * LDA + kk * 4 + k];
                        acc[ct] = mf_mma<MODE>(av, b, acc[ct]);
                    }
                }
            }
        } else {
            const int kg = (lane >> 4) * 8, cl = lane & 15;
#pragma unroll
            for (int kk = 0; kk < KS / 32; kk++) {
                mf_bf8 ah = *(const mf_bf8 *)(Ash + (wave * 16 + cl) * LDA + kk * 32 + kg);
                mf_bf8 al = *(const mf_bf8 *)(Asl + (wave * 16 + cl) * LDA + kk * 32 + kg);
#pragma unroll
                for (int ct = 0; ct < MB_CT; ct++) {
                    if (ct < CT) {
                        mf_bf8 bh = *(const mf_bf8 *)(Bsh + (ct * 16 + cl) * LDA + kk * 32 + kg);
                        mf_bf8 bl = *(const mf_bf8 *)(Bsl + (ct * 16 + cl) * LDA + kk * 32 + kg);
                        acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[ct], 0, 0, 0);
                        acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[ct], 0, 0, 0);
                        acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[ct], 0, 0, 0);
                    }
                }
            }
        }
    }
    // epilogue: Z_m = ex[nx] T_m on the even (real-part) columns, moments,
    // then the sum over the 8 even lanes that share this lane's ions
    double H[4][3], Mm[4][6];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        H[i][0] = H[i][1] = H[i][2] = 0.0;
#pragma unroll
        for (int c = 0; c < 6; c++) Mm[i][c] = 0.0;
    }
    const int cl = lane & 15;
#pragma unroll
    for (int ct = 0; ct < MB_CT; ct++) {
        if (ct >= CT) continue;
        int col = ct * 16 + cl;
        int nx = col / (2 * a.CH), m = (col >> 1) % a.CH;
        bool even = (cl & 1) == 0 && col < a.NC;
        double fnx = (double)nx;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            double tre = (double)acc[ct][i];
            double tim = __shfl_xor(tre, 1, 64);
            int r = (MODE == KM_F64) ? (lane >> 4) + 4 * i : 4 * (lane >> 4) + i;
            int j = jrow0 + wave * 16 + r;
            if (!even || j >= a.nsel) continue;
            size_t ox = 2 * ((size_t)nx * a.nsel + j);
            double xr = (double)tx[ox], xi = (double)tx[ox + 1];
            double zr = xr * tre - xi * tim, zi = xr * tim + xi * tre;
            switch (m) {
            case 0: H[i][0] += fnx * zi; Mm[i][0] += fnx * fnx * zr; break;
            case 1: H[i][1] += zi; Mm[i][1] += fnx * zr; break;
            case 2: H[i][2] += zi; Mm[i][2] += fnx * zr; break;
            case 3: Mm[i][3] += zr; break;
            case 4: Mm[i][4] += zr; break;
            default: Mm[i][5] += zr; break;
            }
        }
    }
    const int nch = (a.CH == 6) ? 9 : 3;
    double *out = a.part + (size_t)blockIdx.y * nch * a.nsel;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        double v[9] = {H[i][0], H[i][1], H[i][2], Mm[i][0], Mm[i][1], Mm[i][2], Mm[i][3], Mm[i][4], Mm[i][5]};
#pragma unroll
        for (int c = 0; c < 9; c++) {
            double s = v[c];
            s += __shfl_xor(s, 2, 64);
            s += __shfl_xor(s, 4, 64);
            s += __shfl_xor(s, 8, 64);
            v[c] = s;
        }
        int r = (MODE == KM_F64) ? (lane >> 4) + 4 * i : 4 * (lane >> 4) + i;
        int j = jrow0 + wave * 16 + r;
        if (cl == 0 && j < a.nsel) {
            for (int c = 0; c < nch; c++) out[(size_t)c * a.nsel + j] = v[c];
        }
    }
}
