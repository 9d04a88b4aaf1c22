// gemm_kernels.h -- the structure-factor sum and its back-projection as
// streaming matrix-core GEMMs over operands stored in HBM.
//
// The ey*ez products of every ion and (ny, nz) point depend only on the
// positions, so they are formed once per force call in two layouts:
//   YZ [p][2j + ri]   (point-major: B operand of the forward sum)
//   YZt[j][2p + ri]   (ion-major:   A operand of the back-projection)
// with K = (ion, re/im) or (point, re/im) contiguous, zero-padded to the
// padded sizes.  Per evaluation only the small operands change:
//   AW[(nx, c, re/im)][2j + ri]   = w_cj ex_j[nx] in the real embedding of
//                                   mfma_kernels.h (rows of the forward sum)
//   WM[(nx, m, re/im)][2p + ri]   = W(nx, p) mom_m(p) (columns of the back sum)
// The big operand is loaded straight into registers, 8 consecutive K values
// per lane; the small one is staged through LDS and shared by the block's
// waves.  For the 16x16x4 FP64/FP32 instructions a lane group g = lane/16
// takes K = 8g + s at step s of an 8-step block, for both operands (any
// permutation of K leaves the sum unchanged); the 16x16x32 FP16
// instruction takes K = 8g..8g+7 in one step.
//
// Split-FP16 (KM_F16X3) keeps 11 + 11 bits of every operand value, but only
// near 1: values below the normal range keep an absolute resolution of
// 2^-24.  The ey*ez products are bounded by 1 already; the weight rows of AW
// (per channel) and the columns of WM (per (nx, m)) are scaled by powers of
// two into [-1, 1] and the sums scaled back afterwards.
#pragma once
#include "mfma_kernels.h"

template <int MODE>
struct GT;
template <>
struct GT<KM_F64> { using T = double; };
template <>
struct GT<KM_F32> { using T = float; };
template <>
struct GT<KM_F16X3> { using T = _Float16; };

// the split mode: two arrays (hi, lo) of 16-bit values and three products
template <int MODE>
constexpr bool g_is_split() { return MODE == KM_F16X3; }

template <int MODE>
using gvec_t = mf_h8;

template <int MODE>
__device__ __forceinline__ mf_f4 g_mma_split(gvec_t<MODE> a, gvec_t<MODE> b, mf_f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

typedef double g_d8 __attribute__((ext_vector_type(8)));
typedef float g_f8 __attribute__((ext_vector_type(8)));

// store x into the operand arrays of a mode (hi and, for split FP16, lo)
template <int MODE>
__device__ __forceinline__ void g_store(void *hi, void *lo, size_t i, double x)
{
    if constexpr (MODE == KM_F64) {
        ((double *)hi)[i] = x;
    } else if constexpr (MODE == KM_F32) {
        ((float *)hi)[i] = (float)x;
    } else {
        // through FP32: a direct double-to-half conversion is a long
        // software sequence; the low part absorbs any double rounding of hi
        _Float16 h = (_Float16)(float)x;
        ((_Float16 *)hi)[i] = h;
        ((_Float16 *)lo)[i] = (_Float16)(float)(x - (double)h);
    }
}

// store the pair (x0, x1) at elements i, i+1 (i even) of a mode's operand
// arrays, one store per array
template <int MODE>
__device__ __forceinline__ void g_store2(void *hi, void *lo, size_t i, double x0, double x1)
{
    if constexpr (MODE == KM_F64) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        *(d2 *)((double *)hi + i) = d2{x0, x1};
    } else if constexpr (MODE == KM_F32) {
        typedef float f2 __attribute__((ext_vector_type(2)));
        *(f2 *)((float *)hi + i) = f2{(float)x0, (float)x1};
    } else {
        // through FP32 (see g_store)
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        const _Float16 a = (_Float16)(float)x0, b = (_Float16)(float)x1;
        *(h2 *)((_Float16 *)hi + i) = h2{a, b};
        *(h2 *)((_Float16 *)lo + i) = h2{(_Float16)(float)(x0 - (double)a), (_Float16)(float)(x1 - (double)b)};
    }
}

// YZ[p][2j + ri] and YZt[j][2p + ri] = (Re, Im)(ey_j[iy] ez_j[iz]),
// p = iy NZ + iz, from one 64-ion x 32-point tile per block: the products go
// through LDS so that both layouts are written in contiguous rows.  Zero
// for j >= nsel or p >= P (padding).  grid (nselp/64, ceil(max(Pp, Pk)/32)).
template <int MODE, typename TT>
__global__ void __launch_bounds__(256) k_yz_build(int nsel, int nselp, int P, int Pp_rows, int Pk, int NZ,
                                                  const TT *__restrict__ ty, const TT *__restrict__ tz, void *yz,
                                                  void *yz_lo, void *yzt, void *yzt_lo)
{
    __shared__ double sr[32][65], si[32][65];   // [point][ion]
    const int j0 = blockIdx.x * 64, p0 = blockIdx.y * 32;
    const int tj = threadIdx.x & 63, tq = threadIdx.x >> 6;
    const int j = j0 + tj;
    for (int pp = tq; pp < 32; pp += 4) {
        const int p = p0 + pp;
        double br = 0.0, bi = 0.0;
        if (j < nsel && p < P) {
            const int iy = p / NZ, iz = p % NZ;
            size_t oy = 2 * ((size_t)iy * nsel + j), oz = 2 * ((size_t)iz * nsel + j);
            double yr = ty[oy], yi = ty[oy + 1], zr = tz[oz], zi = tz[oz + 1];
            br = yr * zr - yi * zi;
            bi = yr * zi + yi * zr;
        }
        sr[pp][tj] = br;
        si[pp][tj] = bi;
    }
    __syncthreads();
    // point-major rows: lanes along ions
    for (int pp = tq; pp < 32; pp += 4) {
        const int p = p0 + pp;
        if (p >= Pp_rows || j >= nselp) continue;
        const size_t o = (size_t)p * 2 * nselp + 2 * (size_t)j;
        g_store2<MODE>(yz, yz_lo, o, sr[pp][tj], si[pp][tj]);
    }
    // ion-major rows: 32 lanes along points, two ions per wavefront pass
    const int tp = threadIdx.x & 31, tr = threadIdx.x >> 5;
    for (int jj = tr; jj < 64; jj += 8) {
        const int jr = j0 + jj, p = p0 + tp;
        if (jr >= nselp || p >= Pk) continue;
        const size_t o = (size_t)jr * 2 * Pk + 2 * (size_t)p;
        g_store2<MODE>(yzt, yzt_lo, o, sr[tp][jj], si[tp][jj]);
    }
}

// the 16 bytes a lane stores per operand array: V values (8 for the 16-bit
// modes, 4 for FP32) of one row, from FP32 values; the split modes write
// hi = round(x) and lo = round(x - hi)
template <int MODE>
struct GPack {
    static constexpr int V = MODE == KM_F32 ? 4 : 8;
    __device__ static void store(void *hi, void *lo, size_t i, const float *x)
    {
        typedef unsigned u4 __attribute__((ext_vector_type(4)));
        if constexpr (MODE == KM_F32) {
            *(u4 *)((float *)hi + i) = u4{__float_as_uint(x[0]), __float_as_uint(x[1]), __float_as_uint(x[2]),
                                          __float_as_uint(x[3])};
        } else {
            using T = typename GT<MODE>::T;
            T h[8], l[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                h[k] = (T)x[k];
                l[k] = (T)(x[k] - (float)h[k]);
            }
            u4 vh, vl;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                vh[k] = (unsigned)__builtin_bit_cast(unsigned short, h[2 * k]) |
                        ((unsigned)__builtin_bit_cast(unsigned short, h[2 * k + 1]) << 16);
                vl[k] = (unsigned)__builtin_bit_cast(unsigned short, l[2 * k]) |
                        ((unsigned)__builtin_bit_cast(unsigned short, l[2 * k + 1]) << 16);
            }
            *(u4 *)((T *)hi + i) = vh;
            *(u4 *)((T *)lo + i) = vl;
        }
    }
};

// k_yz_build for the FP32 and split modes: one 64-ion x 64-point tile per
// block, the products rounded to FP32 in LDS (below the resolution of these
// modes), written in 16-byte stores: rows of 64 ions (point-major) and,
// when yzt is given, of 64 points (ion-major).  nselp is a multiple of 64, Pk of 32
// and Pp_rows of 64.  grid (nselp/64, ceil(max(Pp_rows, Pk)/64)).
template <int MODE>
__global__ void __launch_bounds__(256) k_yz_build16(int nsel, int nselp, int P, int Pp_rows, int Pk, int NZ,
                                                    const float *__restrict__ ty, const float *__restrict__ tz,
                                                    void *yz, void *yz_lo, void *yzt, void *yzt_lo)
{
    __shared__ float sr[64][65], si[64][65];   // [point][ion]
    const int j0 = blockIdx.x * 64, p0 = blockIdx.y * 64;
    {
        const int tj = threadIdx.x & 63, tq = threadIdx.x >> 6;
        const int j = j0 + tj;
        for (int pp = tq; pp < 64; pp += 4) {
            const int p = p0 + pp;
            double br = 0.0, bi = 0.0;
            if (j < nsel && p < P) {
                const int iy = p / NZ, iz = p % NZ;
                const size_t oy = 2 * ((size_t)iy * nsel + j), oz = 2 * ((size_t)iz * nsel + j);
                const double yr = ty[oy], yi = ty[oy + 1], zr = tz[oz], zi = tz[oz + 1];
                br = yr * zr - yi * zi;
                bi = yr * zi + yi * zr;
            }
            sr[pp][tj] = (float)br;
            si[pp][tj] = (float)bi;
        }
    }
    __syncthreads();
    constexpr int V = GPack<MODE>::V;   // values per lane store
    constexpr int E = V / 2;            // complex entries per lane store
    constexpr int LPR = 64 / E;         // lanes per 64-entry row
    constexpr int RPP = 256 / LPR;      // rows per pass of the block
    const int q = threadIdx.x % LPR, rsub = threadIdx.x / LPR;
    float x[V];
    // point-major rows YZ[p][2j + ri]: lanes along ions
    for (int pp = rsub; pp < 64; pp += RPP) {
        const int p = p0 + pp;
        if (p >= Pp_rows) break;
#pragma unroll
        for (int e = 0; e < E; e++) {
            x[2 * e] = sr[pp][q * E + e];
            x[2 * e + 1] = si[pp][q * E + e];
        }
        GPack<MODE>::store(yz, yz_lo, (size_t)p * 2 * nselp + 2 * (size_t)(j0 + q * E), x);
    }
    // ion-major rows YZt[j][2p + ri]: lanes along points (when yzt is given)
    const int pq = p0 + q * E;
    if (yzt && pq < Pk) {
        for (int jj = rsub; jj < 64; jj += RPP) {
#pragma unroll
            for (int e = 0; e < E; e++) {
                x[2 * e] = sr[q * E + e][jj];
                x[2 * e + 1] = si[q * E + e][jj];
            }
            GPack<MODE>::store(yzt, yzt_lo, (size_t)(j0 + jj) * 2 * Pk + 2 * (size_t)pq, x);
        }
    }
}

// AW rows (nx, c, re/im) over K = (j, ri): row (nx,c,0) = (w exr, -w exi),
// row (nx,c,1) = (w exi, w exr).  With the per-block channel maxima wpart
// of k_w_max (nwb blocks), w is scaled by a power of two 2^-e into [-1, 1]
// per channel, and block (0, 0) writes 2^e to winv[c].  grid (nselp/256,
// Kx+1), 256 threads (one wave per channel, C <= 4).  Rows >= M are never
// written (zeroed at allocation).
template <int MODE, typename TT>
__global__ void __launch_bounds__(256) k_aw_build(int nsel, int nselp, int C, const TT *__restrict__ tx,
                                                  const int *__restrict__ idx, const double *__restrict__ q,
                                                  const double *__restrict__ mu, int n, void *aw, void *aw_lo,
                                                  const double *__restrict__ wpart, int nwb, double *winv,
                                                  const int *active)
{
    if (active && !*active) return;
    __shared__ double wsc[4];
    if (wpart) {
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (wv < C) {
            double m = 0.0;
            for (int l = lane; l < nwb; l += 64) m = fmax(m, wpart[(size_t)wv * nwb + l]);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
            if (lane == 0) {
                const int e = m > 0.0 ? ilogb(m) + 1 : 0;
                wsc[wv] = ldexp(1.0, -e);
                if (blockIdx.x == 0 && blockIdx.y == 0) winv[wv] = ldexp(1.0, e);
            }
        }
        __syncthreads();
    }
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = blockIdx.y;
    if (j >= nselp) return;
    double xr = 0.0, xi = 0.0;
    int ion = 0;
    if (j < nsel) {
        size_t ox = 2 * ((size_t)nx * nsel + j);
        xr = tx[ox];
        xi = tx[ox + 1];
        ion = idx ? idx[j] : j;
    }
    const size_t ld = 2 * (size_t)nselp;
    for (int c = 0; c < C; c++) {
        double w = 0.0;
        if (j < nsel) {
            if (C == 1) w = q[ion];
            else if (C == 3) w = mu[(size_t)c * n + ion];
            else w = c == 0 ? q[ion] : mu[(size_t)(c - 1) * n + ion];
            if (wpart) w *= wsc[c];
        }
        const size_t r0 = (size_t)(nx * C + c) * 2;
        g_store2<MODE>(aw, aw_lo, r0 * ld + 2 * (size_t)j, w * xr, -w * xi);
        g_store2<MODE>(aw, aw_lo, (r0 + 1) * ld + 2 * (size_t)j, w * xi, w * xr);
    }
}

// k_aw_build for the FP32 and split modes: E = GPack::V/2 consecutive ions
// per thread (one 16-byte store per row and operand array), one channel
// per block.  grid (ceil(nselp / (256 E)), Kx+1, C), 256 threads.
template <int MODE>
__global__ void __launch_bounds__(256) k_aw_build16(int nsel, int nselp, int C, const float *__restrict__ tx,
                                                    const int *__restrict__ idx, const double *__restrict__ q,
                                                    const double *__restrict__ mu, int n, void *aw, void *aw_lo,
                                                    const double *__restrict__ wpart, int nwb, double *winv,
                                                    const int *active)
{
    if (active && !*active) return;
    constexpr int V = GPack<MODE>::V, E = V / 2;
    __shared__ double wsc[4];
    const int j0 = (blockIdx.x * 256 + threadIdx.x) * E;
    const int nx = blockIdx.y, c = blockIdx.z;
    // the table values and weights are loaded before the channel scales are
    // reduced (the two load chains overlap)
    double xr[E], xi[E], wt[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int j = j0 + e;
        xr[e] = 0.0;
        xi[e] = 0.0;
        wt[e] = 0.0;
        if (j < nsel) {
            const size_t ox = 2 * ((size_t)nx * nsel + j);
            xr[e] = tx[ox];
            xi[e] = tx[ox + 1];
            const int i = idx ? idx[j] : j;
            if (C == 1) wt[e] = q[i];
            else if (C == 3) wt[e] = mu[(size_t)c * n + i];
            else wt[e] = c == 0 ? q[i] : mu[(size_t)(c - 1) * n + i];
        }
    }
    if (wpart) {
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (wv < C) {
            double m = 0.0;
            for (int l = lane; l < nwb; l += 64) m = fmax(m, wpart[(size_t)wv * nwb + l]);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
            if (lane == 0) {
                const int e = m > 0.0 ? ilogb(m) + 1 : 0;
                wsc[wv] = ldexp(1.0, -e);
                if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0) winv[wv] = ldexp(1.0, e);
            }
        }
        __syncthreads();
    }
    if (j0 >= nselp) return;
    const size_t ld = 2 * (size_t)nselp;
    {
        float r0v[V], r1v[V];
#pragma unroll
        for (int e = 0; e < E; e++) {
            const double w = wpart ? wt[e] * wsc[c] : wt[e];
            r0v[2 * e] = (float)(w * xr[e]);
            r0v[2 * e + 1] = (float)(-w * xi[e]);
            r1v[2 * e] = (float)(w * xi[e]);
            r1v[2 * e + 1] = (float)(w * xr[e]);
        }
        const size_t r0 = (size_t)(nx * C + c) * 2;
        GPack<MODE>::store(aw, aw_lo, r0 * ld + 2 * (size_t)j0, r0v);
        GPack<MODE>::store(aw, aw_lo, (r0 + 1) * ld + 2 * (size_t)j0, r1v);
    }
}

// WM columns (nx, m, re/im) over K = (p, ri): with Wm = W(nx, p) mom_m(p),
// column (nx,m,0) = (Re Wm, -Im Wm), column (nx,m,1) = (Im Wm, Re Wm).
// grid (Pp/256, (Kx+1) CH), one thread per point.
template <int MODE>
__global__ void k_wm_build(int P, int Pp, int Ky, int Kz, int CH, const double *__restrict__ W, void *wm,
                           void *wm_lo, const int *active)
{
    if (active && !*active) return;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = blockIdx.y / CH, m = blockIdx.y % CH;
    if (p >= Pp) return;
    double wr = 0.0, wi = 0.0;
    if (p < P) {
        const int NZ = 2 * Kz + 1;
        const int iy = p / NZ, iz = p % NZ;
        const double mom = mb_moment(m, iy - Ky, iz - Kz);
        const size_t kf = (size_t)nx * P + p;
        wr = W[2 * kf] * mom;
        wi = W[2 * kf + 1] * mom;
    }
    const size_t ld = 2 * (size_t)Pp;
    const size_t c0 = (size_t)(nx * CH + m) * 2;
    g_store2<MODE>(wm, wm_lo, c0 * ld + 2 * (size_t)p, wr, -wi);
    g_store2<MODE>(wm, wm_lo, (c0 + 1) * ld + 2 * (size_t)p, wi, wr);
}

// k_sk_reduce and the unscaled WM columns of k_wm_build in one pass, one
// thread per k point (nx, p): W(k), the energy and stress terms, then the
// columns (nx, m, re/im) at point p for the CH moments
template <int MODE>
__global__ void k_sk_reduce_wm(KRedArgs a, int CH, int Pk, void *wm, void *wm_lo)
{
    if (a.active && !*a.active) return;
    const size_t nkf = (size_t)(a.Kx + 1) * a.P;
    const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (k >= nkf) return;
    double Wr, Wi;
    sk_reduce_one(a, k, Wr, Wi);
    const int nx = (int)(k / a.P), p = (int)(k % a.P);
    const int NZ = 2 * a.Kz + 1;
    const int iy = p / NZ, iz = p % NZ;
    const size_t ld = 2 * (size_t)Pk;
    for (int m = 0; m < CH; m++) {
        const double mom = mb_moment(m, iy - a.Ky, iz - a.Kz);
        const double wr = Wr * mom, wi = Wi * mom;
        const size_t c0 = (size_t)(nx * CH + m) * 2;
        g_store2<MODE>(wm, wm_lo, c0 * ld + 2 * (size_t)p, wr, -wi);
        g_store2<MODE>(wm, wm_lo, (c0 + 1) * ld + 2 * (size_t)p, wi, wr);
    }
}

// the reduction (sk_reduce_one) of a split-FP16 evaluation, one thread per
// k-vector, with the column maxima of its WM columns: for every nx plane in
// the block and moment m < CH, the block maximum of |W mom_m| goes into
// cm[nx CH + m] by atomicMax on the bits of the non-negative double (cm
// zeroed by k_part_sum beforehand; non-finite values count as zero)
template <int CH>
__global__ void __launch_bounds__(256) k_sk_reduce_cm(KRedArgs a, unsigned long long *cm)
{
    if (a.active && !*a.active) return;
    __shared__ double red[4];
    const size_t nkf = (size_t)(a.Kx + 1) * a.P;
    const size_t k0 = (size_t)blockIdx.x * 256, k = k0 + threadIdx.x;
    const int NZ = 2 * a.Kz + 1;
    double v[CH];
#pragma unroll
    for (int m = 0; m < CH; m++) v[m] = 0.0;
    int nx = -1;
    if (k < nkf) {
        double Wr, Wi;
        sk_reduce_one(a, k, Wr, Wi);
        nx = (int)(k / a.P);
        const int p = (int)(k % a.P), iy = p / NZ, iz = p % NZ;
#pragma unroll
        for (int m = 0; m < CH; m++) {
            const double mom = mb_moment(m, iy - a.Ky, iz - a.Kz);
            const double x = fmax(fabs(Wr * mom), fabs(Wi * mom));
            v[m] = isfinite(x) ? x : 0.0;
        }
    }
    const size_t kl = k0 + 255 < nkf ? k0 + 255 : nkf - 1;
    const int nlo = (int)(k0 / a.P), nhi = (int)(kl / a.P);
    for (int nxv = nlo; nxv <= nhi; nxv++) {
#pragma unroll
        for (int m = 0; m < CH; m++) {
            double x = nx == nxv ? v[m] : 0.0;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o));
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
            __syncthreads();
            if (threadIdx.x == 0) {
                const double b = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
                if (b > 0.0) atomicMax(cm + nxv * CH + m, (unsigned long long)__double_as_longlong(b));
            }
            __syncthreads();
        }
    }
}

// k_part_sum and k_sk_reduce_cm in one pass over the forward sum's FP32
// partial sums part[nchunk][M][P] (rows (nx, c, re/im), M = (Kx+1) 2C):
// PSUM_PB k-vectors per block, PSUM_G thread groups each summing every
// PSUM_G-th K range of the 2C rows of a k-vector (loads coalesced over the
// points), the groups added in a fixed order, channel c times inv[c]; then
// the reduction of sk_reduce_S and the column maxima of k_sk_reduce_cm (cm
// zeroed by the forward sum's block (0, 0)).  grid ceil(nkf / PSUM_PB).
#define PSUM_PB 64
#define PSUM_G 4
#define PSUM_U 2   // K ranges a thread loads per pass
template <int C, int CH>
__global__ void __launch_bounds__(256) k_psum_reduce_cm(KRedArgs a, const float *__restrict__ part, int nchunk,
                                                        int M, const double *__restrict__ inv,
                                                        unsigned long long *cm)
{
    static_assert(PSUM_PB * PSUM_G == 256 && PSUM_PB <= 64, "block shape");
    if (a.active && !*a.active) return;
    constexpr int R = 2 * C;   // rows per k-vector
    __shared__ double sh[PSUM_G - 1][R][PSUM_PB];
    const int P = a.P;
    const int nkf = (a.Kx + 1) * P;
    const int k0 = blockIdx.x * PSUM_PB;
    const int l = threadIdx.x % PSUM_PB, g = threadIdx.x / PSUM_PB;
    const int k = k0 + l;
    const size_t nval = (size_t)M * P;
    double s[R];
#pragma unroll
    for (int r = 0; r < R; r++) s[r] = 0.0;
    const bool live = k < nkf && a.Ak[k] != 0.0;
    if (live) {
        const int nx = k / P, p = k - nx * P;
        const float *pb = part + (size_t)nx * R * P + p;
        // K ranges g, g + G, ..., PSUM_U at a time (all row loads issued
        // first), then one at a time
        int ch = g;
        for (; ch + (PSUM_U - 1) * PSUM_G < nchunk; ch += PSUM_U * PSUM_G) {
            float x[PSUM_U][R];
#pragma unroll
            for (int u = 0; u < PSUM_U; u++)
#pragma unroll
                for (int r = 0; r < R; r++) x[u][r] = pb[(size_t)(ch + u * PSUM_G) * nval + (size_t)r * P];
#pragma unroll
            for (int r = 0; r < R; r++) {
                double t = 0.0;
#pragma unroll
                for (int u = 0; u < PSUM_U; u++) t += (double)x[u][r];
                s[r] += t;
            }
        }
        for (; ch < nchunk; ch += PSUM_G) {
            float x0[R];
#pragma unroll
            for (int r = 0; r < R; r++) x0[r] = pb[(size_t)ch * nval + (size_t)r * P];
#pragma unroll
            for (int r = 0; r < R; r++) s[r] += (double)x0[r];
        }
    }
    if (g > 0) {
#pragma unroll
        for (int r = 0; r < R; r++) sh[g - 1][r][l] = s[r];
    }
    __syncthreads();
    // wave 0 finishes: its lanes of group 0 own the block's k-vectors, the
    // others take part in the wave maxima with nothing
    if (threadIdx.x >= 64) return;
    const bool own = g == 0 && k < nkf;
    double v[CH];
#pragma unroll
    for (int m = 0; m < CH; m++) v[m] = 0.0;
    int nx = -1;
    if (own) {
        double S[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
        if (live) {
#pragma unroll
            for (int c = 0; c < C; c++) {
                const double sc = inv ? inv[c] : 1.0;
#pragma unroll
                for (int ri = 0; ri < 2; ri++) {
                    const int r = 2 * c + ri;
                    double t = s[r];
#pragma unroll
                    for (int q = 0; q < PSUM_G - 1; q++) t += sh[q][r][l];
                    S[c][ri] = t * sc;
                }
            }
        }
        double Wr, Wi;
        sk_reduce_S(a, (size_t)k, S, Wr, Wi);
        nx = k / P;
        const int NZ = 2 * a.Kz + 1;
        const int p = k - nx * P, iy = p / NZ, iz = p % NZ;
#pragma unroll
        for (int m = 0; m < CH; m++) {
            const double mom = mb_moment(m, iy - a.Ky, iz - a.Kz);
            const double x = fmax(fabs(Wr * mom), fabs(Wi * mom));
            v[m] = isfinite(x) ? x : 0.0;
        }
    }
    // block maxima per nx plane of the block's k-vectors (one wave)
    const int kl = k0 + PSUM_PB - 1 < nkf ? k0 + PSUM_PB - 1 : nkf - 1;
    const int nlo = k0 / P, nhi = kl / P;
    for (int nxv = nlo; nxv <= nhi; nxv++) {
#pragma unroll
        for (int m = 0; m < CH; m++) {
            double x = nx == nxv ? v[m] : 0.0;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o));
            if (threadIdx.x == 0 && x > 0.0) atomicMax(cm + nxv * CH + m, (unsigned long long)__double_as_longlong(x));
        }
    }
}

// split-FP16 WM columns (nx, m, re/im) over the Pk points (zero past P),
// each column scaled by 2^-e into [-1, 1] from its maximum cm[nx CH + m],
// 2^e into cinv.  grid (Pk/256, (Kx+1) CH).
__global__ void __launch_bounds__(256) k_wm_scaled(int P, int Pk, int Ky, int Kz, int CH, const double *__restrict__ W,
                                                   const unsigned long long *__restrict__ cm, void *wm, void *wm_lo,
                                                   double *cinv, const int *active)
{
    if (active && !*active) return;
    const int nx = blockIdx.y / CH, m = blockIdx.y % CH;
    const double amax = __longlong_as_double((long long)cm[blockIdx.y]);
    const int e = amax > 0.0 ? ilogb(amax) + 1 : 0;
    const double scl = ldexp(1.0, -e);
    if (blockIdx.x == 0 && threadIdx.x == 0) cinv[blockIdx.y] = ldexp(1.0, e);
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= Pk) return;
    const int NZ = 2 * Kz + 1;
    double wr = 0.0, wi = 0.0;
    if (p < P) {
        const int iy = p / NZ, iz = p % NZ;
        const double mom = mb_moment(m, iy - Ky, iz - Kz) * scl;
        const size_t kf = (size_t)nx * P + p;
        wr = W[2 * kf] * mom;
        wi = W[2 * kf + 1] * mom;
    }
    const size_t ld = 2 * (size_t)Pk;
    const size_t c0 = (size_t)(nx * CH + m) * 2;
    g_store2<KM_F16X3>(wm, wm_lo, c0 * ld + 2 * (size_t)p, wr, -wi);
    g_store2<KM_F16X3>(wm, wm_lo, (c0 + 1) * ld + 2 * (size_t)p, wi, wr);
}

// largest |w| of each weight channel (q, mu components) over the block's
// 256 selected ions: wpart[c nb + block], nb = gridDim.x.  Non-finite
// weights count as zero.
__global__ void __launch_bounds__(256) k_w_max(int nsel, int C, const int *__restrict__ idx,
                                               const double *__restrict__ q, const double *__restrict__ mu, int n,
                                               double *__restrict__ wpart, const int *active)
{
    if (active && !*active) return;
    __shared__ double red[4][4];   // [wave][channel]
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int ion = (j < nsel) ? (idx ? idx[j] : j) : 0;
    // all channels' loads first, then one wave reduction per channel
    double w[4] = {0.0, 0.0, 0.0, 0.0};
    if (j < nsel) {
        if (C == 1) {
            w[0] = q[ion];
        } else if (C == 3) {
#pragma unroll
            for (int c = 0; c < 3; c++) w[c] = mu[(size_t)c * n + ion];
        } else {
            w[0] = q[ion];
#pragma unroll
            for (int c = 0; c < 3; c++) w[c + 1] = mu[(size_t)c * n + ion];
        }
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
        double x = isfinite(w[c]) ? fabs(w[c]) : 0.0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][c] = x;
    }
    __syncthreads();
    if ((int)threadIdx.x < C)
        wpart[(size_t)threadIdx.x * gridDim.x + blockIdx.x] =
            fmax(fmax(red[0][threadIdx.x], red[1][threadIdx.x]), fmax(red[2][threadIdx.x], red[3][threadIdx.x]));
}

// WM columns as k_wm_build, each (nx, m) pair scaled by a power of two
// into [-1, 1]; cinv[nx CH + m] receives the inverse scale.  One block per
// (nx, m): grid (Kx+1) CH.
template <int MODE>
__global__ void __launch_bounds__(256) k_wm_build_scaled(int P, int Pp, int Ky, int Kz, int CH,
                                                         const double *__restrict__ W, void *wm, void *wm_lo,
                                                         double *cinv, const int *active)
{
    if (active && !*active) return;
    __shared__ double red[256];
    const int nx = blockIdx.x / CH, m = blockIdx.x % CH;
    const int NZ = 2 * Kz + 1;
    double mx = 0.0;
    for (int p = threadIdx.x; p < P; p += 256) {
        const int iy = p / NZ, iz = p % NZ;
        const double mom = mb_moment(m, iy - Ky, iz - Kz);
        const size_t kf = (size_t)nx * P + p;
        mx = fmax(mx, fmax(fabs(W[2 * kf] * mom), fabs(W[2 * kf + 1] * mom)));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    const double amax = red[0];
    const int e = (amax > 0.0 && isfinite(amax)) ? ilogb(amax) + 1 : 0;
    const double scl = ldexp(1.0, -e);
    if (threadIdx.x == 0) cinv[blockIdx.x] = ldexp(1.0, e);
    const size_t ld = 2 * (size_t)Pp;
    const size_t c0 = (size_t)(nx * CH + m) * 2;
    for (int p = threadIdx.x; p < Pp; p += 256) {
        double wr = 0.0, wi = 0.0;
        if (p < P) {
            const int iy = p / NZ, iz = p % NZ;
            const double mom = mb_moment(m, iy - Ky, iz - Kz) * scl;
            const size_t kf = (size_t)nx * P + p;
            wr = W[2 * kf] * mom;
            wi = W[2 * kf + 1] * mom;
        }
        g_store2<MODE>(wm, wm_lo, c0 * ld + 2 * (size_t)p, wr, -wi);
        g_store2<MODE>(wm, wm_lo, (c0 + 1) * ld + 2 * (size_t)p, wi, wr);
    }
}

// 8 consecutive operand values of one lane
template <int MODE>
struct GRun {
    using T = typename GT<MODE>::T;
    typename std::conditional<MODE == KM_F64, g_d8,
                              typename std::conditional<MODE == KM_F32, g_f8, gvec_t<MODE>>::type>::type v;
};

template <int MODE>
__device__ __forceinline__ GRun<MODE> g_load(const void *base, size_t off)
{
    GRun<MODE> r;
    if constexpr (MODE == KM_F64) r.v = *(const g_d8 *)((const double *)base + off);
    else if constexpr (MODE == KM_F32) r.v = *(const g_f8 *)((const float *)base + off);
    else r.v = *(const mf_h8 *)((const _Float16 *)base + off);
    return r;
}

// the run of 8 operand values (re, im of 4 entries y_i z_i) formed from
// FP32 table values: products in FP32, then for the split modes
// hi = round(x) and lo = round(x - hi)
template <int MODE>
__device__ __forceinline__ void g_gen_run(const float2 *y, const float2 *z, GRun<MODE> &hi, GRun<MODE> &lo)
{
    float x[8];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        x[2 * i] = fmaf(y[i].x, z[i].x, -(y[i].y * z[i].y));
        x[2 * i + 1] = fmaf(y[i].x, z[i].y, y[i].y * z[i].x);
    }
    if constexpr (MODE == KM_F32) {
#pragma unroll
        for (int k = 0; k < 8; k++) hi.v[k] = x[k];
    } else {
        using T = typename GT<MODE>::T;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const T h = (T)x[k];
            hi.v[k] = h;
            lo.v[k] = (T)(x[k] - (float)h);
        }
    }
}

#define GK 32          // K values per block step
#define GLDS_PAD 8     // LDS row padding (elements)

struct GFwdArgs {
    const void *aw, *aw_lo;   // [16 MF_MT][ldk], rows >= M zero
    const void *yz, *yz_lo;   // [Pp][ldk] (FP64 mode)
    // the other modes form YZ[p][2j + ri] = ey_j[iy] ez_j[iz] from the FP32
    // tables [NY][nsel][2], [NZ][nsel][2] of the selected ions
    const float *gty, *gtz;
    int nsel, NY, NZ;
    int MT, M, P, C, ldk, kchunk;
    double *part;             // [nchunk][M][P] (FP64, or FP32 in the other modes)
    const int *active;
    unsigned long long *cmz;  // when set, block (0, 0) zeroes these ncmz column maxima
    int ncmz;                 // for the k_psum_reduce_cm that follows
    // comb (three dipole channels, rows (nx, c, re/im) with M = 6 (Kx+1)):
    // the partial sums are written as rho = i k.S_mu per box point, rows
    // (nx, re/im), with the channels' weight scales cinv[3] and the k vectors
    // kvec[(nx, p)][3] folded in (the reduction that follows takes them as
    // one charge-like channel)
    int comb;
    const double *kvec, *cinv;
};

// LDS of a forward launch with the combined epilogue: the 4 waves' tiles
inline size_t g_fwd_comb_lds_bytes(int MT, int CW) { return (size_t)4 * MT * 16 * (CW * 16 + 1) * sizeof(float); }

// row tiles a forward kernel is instantiated for: every launch computes
// this many (zero rows past M) so that no matrix instruction is predicated
inline int g_mt_inst(int MT) { return MT <= 3 ? 3 : (MT <= 5 ? 5 : (MT <= 7 ? 7 : 9)); }

// table slice of a forward step formed from the tables (GEN): GT_ROWS rows
// (the ey rows of the block's points, then all ez rows) x 16 ions, rows of
// GT_LD complex values (padded: the 16 rows a fragment reads fall in
// distinct banks)
#define GT_ROWS 64
#define GT_LD 18

template <int MODE>
inline size_t g_fwd_lds_bytes(int MT, bool gen = false)
{
    size_t e = (size_t)MT * 16 * (GK + GLDS_PAD);
    e = MODE == KM_F64 ? e * 8 : (MODE == KM_F32 ? e * 4 : e * 2 * 2);
    return e + (gen ? (size_t)GT_ROWS * GT_LD * sizeof(float2) : 0);
}

// whether a forward block's table slice fits GT_ROWS rows: at most
// (64 CW - 1)/NZ + 2 ey rows and NZ ez rows
inline bool g_fwd_gen_fits(int CW, int NZ) { return (64 * CW - 1) / NZ + 2 + NZ <= GT_ROWS; }

// forward sum: block = 4 waves x CW column tiles of 16 points, all MT row
// tiles, one K range (grid.y); partial sums per K range in the layout of
// k_sk_reduce.  The next step's A rows and B runs are loaded while the
// current step computes; consecutive instructions update different
// accumulators (no back-to-back dependence on one accumulator).
#define GF_SLOTS 3   // A runs a thread stages per step: ceil(16 MT GK/8 / 256) for MT <= 12

// Block order of the forward sum: consecutive workgroups are dispatched to
// the 8 XCDs in turn, and blocks of one K range share that range of AW.
// The dispatch index w is mapped so that each XCD receives a contiguous run
// of (x fastest) logical blocks, so a K range is fetched into one or two
// XCD L2s instead of all eight.  A bijection: correctness does not depend
// on where blocks run.
__device__ __forceinline__ void g_xcd_order(int gx, int gy, int &bx, int &by)
{
    const int w = blockIdx.x + blockIdx.y * gx;
    const int total = gx * gy;
    const int q = total >> 3, r = total & 7;
    const int xcd = w & 7, i = w >> 3;
    const int L = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + i;
    bx = L % gx;
    by = L / gx;
}
template <int MODE, int CW, int MTMAX, bool GEN = false>
__global__ void __launch_bounds__(256) k_gemm_fwd(GFwdArgs a)
{
    static_assert(!GEN || MODE != KM_F64, "the FP64 mode loads stored operands");
    if (a.active && !*a.active) return;
    if (a.cmz && blockIdx.x == 0 && blockIdx.y == 0)
        for (int i = threadIdx.x; i < a.ncmz; i += 256) a.cmz[i] = 0ull;
    using T = typename GT<MODE>::T;
    constexpr int LD = GK + GLDS_PAD;
    constexpr bool SPLIT = g_is_split<MODE>();
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *As = (T *)smem_raw;
    T *Asl = As + MTMAX * 16 * LD;   // low parts of the split modes
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, cl = lane & 15;
    int bx, by;
    g_xcd_order(gridDim.x, gridDim.y, bx, by);
    const int ct0 = (bx * 4 + wave) * CW;
    const int k0 = by * a.kchunk;
    const int k1 = min(k0 + a.kchunk, a.ldk);
    constexpr int nslot = MTMAX * 16 * (GK / 8);
    static_assert(nslot <= 256 * GF_SLOTS, "A staging slots");
    using acc_t = typename std::conditional<MODE == KM_F64, mf_d4, mf_f4>::type;
    acc_t acc[MTMAX][CW];
#pragma unroll
    for (int t = 0; t < MTMAX; t++)
#pragma unroll
        for (int c = 0; c < CW; c++) acc[t][c] = acc_t{0, 0, 0, 0};
    // the B runs: loaded (stored operands) or, with GEN, formed from a table
    // slice staged in LDS per step: rows 0..nyr-1 the ey rows iy0.. of the
    // block's points, then the NZ ez rows, each over the step's 16 ions
    // (points past P use the last point, ions past nsel the last ion,
    // against zero AW columns)
    size_t brow[CW];
#pragma unroll
    for (int c = 0; c < CW; c++) brow[c] = (size_t)((ct0 + c) * 16 + cl) * a.ldk;
    float2 *ts = (float2 *)(smem_raw + (size_t)MTMAX * 16 * LD * sizeof(T) * (SPLIT ? 2 : 1));
    int iy0 = 0, nyr = 0, nrow = 0;
    int yr[CW], zr[CW];
    if constexpr (GEN) {
        const int pb0 = bx * 4 * CW * 16;
        const int pl = min(pb0 + 64 * CW - 1, a.P - 1);
        iy0 = min(pb0, a.P - 1) / a.NZ;
        nyr = pl / a.NZ - iy0 + 1;
        nrow = nyr + a.NZ;
#pragma unroll
        for (int c = 0; c < CW; c++) {
            const int p = min((ct0 + c) * 16 + cl, a.P - 1);
            const int iy = p / a.NZ;
            yr[c] = (iy - iy0) * GT_LD + 4 * g;
            zr[c] = (nyr + p - iy * a.NZ) * GT_LD + 4 * g;
        }
    }
    // Two register sets alternate (the K range holds an even number of
    // steps): the loads of step k+2 are issued before step k computes, and
    // never under a branch, so the compiler keeps them in flight across the
    // step (a conditional load makes it wait for everything before the
    // matrix instructions).  Past the end the last step is loaded again.
    constexpr int NQ = (nslot + 255) / 256;
    constexpr int TSL = GT_ROWS * 16 / 256;   // table entries a thread stages per step
    struct SetS {
        GRun<MODE> bh[CW], bl[CW], ra[NQ], rl[NQ];
    };
    struct SetG {
        float2 t[TSL];
        GRun<MODE> ra[NQ], rl[NQ];
    };
    using Set = typename std::conditional<GEN, SetG, SetS>::type;
    Set S0, S1;
    auto load_step = [&](Set &S, int kb) {
        if constexpr (GEN) {
            // entry e = (row, ion): rows past nrow repeat the last row
#pragma unroll
            for (int u = 0; u < TSL; u++) {
                const int e = threadIdx.x + 256 * u;
                const int row = min(e >> 4, nrow - 1), jj = min((kb >> 1) + (e & 15), a.nsel - 1);
                const float *src = row < nyr ? a.gty + 2 * ((size_t)(iy0 + row) * a.nsel + jj)
                                             : a.gtz + 2 * ((size_t)(row - nyr) * a.nsel + jj);
                S.t[u] = *(const float2 *)src;
            }
        } else {
#pragma unroll
            for (int c = 0; c < CW; c++) {
                S.bh[c] = g_load<MODE>(a.yz, brow[c] + kb + 8 * g);
                if constexpr (SPLIT) S.bl[c] = g_load<MODE>(a.yz_lo, brow[c] + kb + 8 * g);
            }
        }
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int e0 = threadIdx.x + 256 * q, e = e0 < nslot ? e0 : nslot - 1;
            const int r = e / (GK / 8), q8 = (e % (GK / 8)) * 8;
            S.ra[q] = g_load<MODE>(a.aw, (size_t)r * a.ldk + kb + q8);
            if constexpr (SPLIT) S.rl[q] = g_load<MODE>(a.aw_lo, (size_t)r * a.ldk + kb + q8);
        }
    };
    auto step = [&](Set &S, int kb) {
        // this step's A rows (and table slice) into LDS; stored B runs stay
        // in registers
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int e = threadIdx.x + 256 * q;
            if (e < nslot) {
                const int r = e / (GK / 8), q8 = (e % (GK / 8)) * 8;
                *(decltype(S.ra[q].v) *)(As + r * LD + q8) = S.ra[q].v;
                if constexpr (SPLIT) *(decltype(S.rl[q].v) *)(Asl + r * LD + q8) = S.rl[q].v;
            }
        }
        GRun<MODE> ch[CW], cl_[CW];
        if constexpr (GEN) {
#pragma unroll
            for (int u = 0; u < TSL; u++) {
                const int e = threadIdx.x + 256 * u;
                if ((e >> 4) < nrow) ts[(e >> 4) * GT_LD + (e & 15)] = S.t[u];
            }
        } else {
#pragma unroll
            for (int c = 0; c < CW; c++) {
                ch[c] = S.bh[c];
                if constexpr (SPLIT) cl_[c] = S.bl[c];
            }
        }
        __syncthreads();
        load_step(S, kb + 2 * GK < k1 ? kb + 2 * GK : kb);
        if constexpr (GEN) {
#pragma unroll
            for (int c = 0; c < CW; c++) {
                typedef float f4 __attribute__((ext_vector_type(4)));
                float2 y[4], z[4];
                const f4 y01 = *(const f4 *)(ts + yr[c]), y23 = *(const f4 *)(ts + yr[c] + 2);
                const f4 z01 = *(const f4 *)(ts + zr[c]), z23 = *(const f4 *)(ts + zr[c] + 2);
                y[0] = float2{y01[0], y01[1]}; y[1] = float2{y01[2], y01[3]};
                y[2] = float2{y23[0], y23[1]}; y[3] = float2{y23[2], y23[3]};
                z[0] = float2{z01[0], z01[1]}; z[1] = float2{z01[2], z01[3]};
                z[2] = float2{z23[0], z23[1]}; z[3] = float2{z23[2], z23[3]};
                g_gen_run<MODE>(y, z, ch[c], cl_[c]);
            }
        }
        if constexpr (SPLIT) {
            using V = gvec_t<MODE>;
            V ah[MTMAX], al[MTMAX];
#pragma unroll
            for (int t = 0; t < MTMAX; t++) {
                ah[t] = *(const V *)(As + (t * 16 + cl) * LD + 8 * g);
                al[t] = *(const V *)(Asl + (t * 16 + cl) * LD + 8 * g);
            }
#pragma unroll
            for (int t = 0; t < MTMAX; t++)
#pragma unroll
                for (int c = 0; c < CW; c++) acc[t][c] = g_mma_split<MODE>(ah[t], ch[c].v, acc[t][c]);
#pragma unroll
            for (int t = 0; t < MTMAX; t++)
#pragma unroll
                for (int c = 0; c < CW; c++) acc[t][c] = g_mma_split<MODE>(ah[t], cl_[c].v, acc[t][c]);
#pragma unroll
            for (int t = 0; t < MTMAX; t++)
#pragma unroll
                for (int c = 0; c < CW; c++) acc[t][c] = g_mma_split<MODE>(al[t], ch[c].v, acc[t][c]);
        } else {
#pragma unroll
            for (int s = 0; s < 8; s++) {
#pragma unroll
                for (int t = 0; t < MTMAX; t++) {
                    const T av = As[(t * 16 + cl) * LD + 8 * g + s];
#pragma unroll
                    for (int c = 0; c < CW; c++) acc[t][c] = mf_mma<MODE>(av, ch[c].v[s], acc[t][c]);
                }
            }
        }
    };
    if (k0 < k1) {
        load_step(S0, k0);
        load_step(S1, k0 + GK < k1 ? k0 + GK : k0);
    }
    for (int kb = k0; kb < k1; kb += 2 * GK) {
        step(S0, kb);
        step(S1, kb + GK);
    }
    if constexpr (MODE != KM_F64) {
        if (a.comb) {
            // rho = i k.S_mu of the wave's points: its tiles through LDS (every
            // wave is past the K loop's last LDS read first), then one (nx,
            // point) per lane and pass
            constexpr int TLD = CW * 16 + 1;
            __syncthreads();
            float *tw = (float *)smem_raw + (size_t)wave * MTMAX * 16 * TLD;
#pragma unroll
            for (int t = 0; t < MTMAX; t++)
#pragma unroll
                for (int c = 0; c < CW; c++)
#pragma unroll
                    for (int i = 0; i < 4; i++) tw[(t * 16 + 4 * g + i) * TLD + c * 16 + cl] = acc[t][c][i];
            __syncthreads();
            const double s0 = a.cinv ? a.cinv[0] : 1.0, s1 = a.cinv ? a.cinv[1] : 1.0, s2 = a.cinv ? a.cinv[2] : 1.0;
            const int KX1 = a.M / 6;
            float *outc = (float *)a.part + (size_t)by * 2 * KX1 * a.P;
            for (int q = lane; q < KX1 * CW * 16; q += 64) {
                const int nx = q / (CW * 16), cc = q - nx * (CW * 16);
                const int p = ct0 * 16 + cc;
                if (p >= a.P) continue;
                const float *r0 = tw + (size_t)(nx * 6) * TLD + cc;
                const size_t kf = (size_t)nx * a.P + p;
                const double kx = a.kvec[3 * kf] * s0, ky = a.kvec[3 * kf + 1] * s1, kz = a.kvec[3 * kf + 2] * s2;
                const double sxr = r0[0], sxi = r0[TLD], syr = r0[2 * TLD], syi = r0[3 * TLD], szr = r0[4 * TLD],
                             szi = r0[5 * TLD];
                outc[(size_t)(2 * nx) * a.P + p] = (float)(-(kx * sxi + ky * syi + kz * szi));
                outc[(size_t)(2 * nx + 1) * a.P + p] = (float)(kx * sxr + ky * syr + kz * szr);
            }
            return;
        }
    }
    // partial sums of this K range, [row][point] (lanes write consecutive
    // points; FP32 accumulators are stored as FP32)
    using part_t = typename std::conditional<MODE == KM_F64, double, float>::type;
    part_t *out = (part_t *)a.part + (size_t)by * a.M * a.P;
#pragma unroll
    for (int t = 0; t < MTMAX; t++) {
#pragma unroll
        for (int c = 0; c < CW; c++) {
            const int p = (ct0 + c) * 16 + cl;
            if (p >= a.P) continue;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = (MODE == KM_F64) ? t * 16 + g + 4 * i : t * 16 + 4 * g + i;
                if (r < a.M) out[(size_t)r * a.P + p] = (part_t)acc[t][c][i];
            }
        }
    }
}

// sum of the K-range partials of the forward sum, one thread per value
// (row r = (nx, c, re/im), point p), chunks in a fixed order (bitwise
// reproducible), into the layout [(nx, p)][c][re/im] of k_sk_reduce; values
// are multiplied by inv[c] when the weights were scaled.  Block 0 also
// zeroes the ncm column maxima cm of the reduction that follows
// (k_sk_reduce_cm), if given.
template <typename part_t>
__global__ void k_part_sum(const part_t *__restrict__ part, int nchunk, int M, int P, double *__restrict__ out,
                           int C, const double *__restrict__ inv, const int *active,
                           unsigned long long *cm = nullptr, int ncm = 0)
{
    if (active && !*active) return;
    if (cm && blockIdx.x == 0)
        for (int i = threadIdx.x; i < ncm; i += blockDim.x) cm[i] = 0ull;
    const size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t nval = (size_t)M * P;
    if (v >= nval) return;
    const int r = (int)(v / P), p = (int)(v % P);
    // loads issued eight at a time, summed in chunk order
    double s = 0.0;
    int ch = 0;
    for (; ch + 8 <= nchunk; ch += 8) {
        part_t x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = part[(size_t)(ch + u) * nval + v];
#pragma unroll
        for (int u = 0; u < 8; u++) s += (double)x[u];
    }
    for (; ch < nchunk; ch++) s += (double)part[(size_t)ch * nval + v];
    const int nx = r / (2 * C), cc = (r >> 1) % C, ri = r & 1;
    if (inv) s *= inv[cc];
    out[((size_t)nx * P + p) * (2 * C) + 2 * cc + ri] = s;
}

#define GB_SLOTS 2   // B runs a thread stages per step: ceil(16 CT GK/8 / 512) for CT <= 16

struct GBackArgs {
    const void *yzt, *yzt_lo;  // [nselp][ldk] (FP64 mode)
    // the other modes form YZt[j][2p + ri] = ey_j[iy] ez_j[iz] from the FP32
    // tables [NY][nsel][2], [NZ][nsel][2] of the selected ions
    const float *gty, *gtz;
    int NY, NZ;
    const void *wm, *wm_lo;    // [16 CT][ldk]
    const void *tx;            // ex tables [Kx+1][nsel][2] (TT)
    const double *csc;         // inverse column scales [(Kx+1) CH] (scaled WM) or null
    int nsel, CH, NC, CT, ldk, kchunk;
    double *part;              // [nsplit][3 or 9][nsel]
    const int *active;
    // grid rows t2_y0.. (none when t2_y0 = gridDim.y): the real-space field
    // of the dipoles t2mu at the polarisable ions from the FP32 pair tensors
    // (rows of k_t2_build), 32 ions per block, into t2out [3][n]
    int t2_y0, t2_npol, t2_capP, t2_n;
    const float *t2;
    const int *t2_cnt, *t2_idx, *t2_pidx;
    const double *t2mu;
    double *t2out;
};

// the pair-tensor blocks of k_gemm_back: 16 lanes per polarisable ion, as
// k_sk_back_fin_t2; every polarisable ion's field is written (zero for an
// ion without dipole)
__device__ __forceinline__ void back_t2_rows(const GBackArgs &a)
{
    constexpr int L = 16;
    const int lane = threadIdx.x % L;
    const int t = (((int)blockIdx.y - a.t2_y0) * (int)gridDim.x + (int)blockIdx.x) * (512 / L) + (int)threadIdx.x / L;
    const int n = a.t2_n;
    const double *mu = a.t2mu;
    float ex = 0, ey = 0, ez = 0;
    int i = -1;
    if (t < a.t2_npol) {
        i = a.t2_pidx[t];
        if ((mu[i] != 0.0) || (mu[n + i] != 0.0) || (mu[2 * n + i] != 0.0))
            t2_row_sum<float, L>((size_t)t * a.t2_capP, a.t2_cnt[t], lane, (size_t)a.t2_npol * a.t2_capP, a.t2_idx,
                                 a.t2, mu, n, ex, ey, ez);
    }
    const double dx = group_sum<L>((double)ex), dy = group_sum<L>((double)ey), dz = group_sum<L>((double)ez);
    if (i >= 0 && lane == 0) {
        a.t2out[i] = dx;
        a.t2out[n + i] = dy;
        a.t2out[2 * n + i] = dz;
    }
}

// column tiles a back-projection kernel is instantiated for (every launch
// computes this many; columns past NC are zero or ignored)
inline int g_ct_inst(int CT) { return CT <= 4 ? 4 : (CT <= 7 ? 7 : (CT <= 13 ? 13 : MB_CT)); }

// largest k box of the back-projection: (Kx+1) 6 2 <= 16 MB_CT
#define GB_KX1_MAX (16 * MB_CT / 12)

// epilogue of the back-projection: waves whose accumulator tiles one round
// stages (all 8 where the LDS allows), threads per ion (each takes every
// TPI-th nx), and the padded row of the block's ex table (elements: the TPI
// lanes of one ion read rows in distinct banks)
constexpr int gb_rw(int mode, int ct) { return mode == KM_F64 ? 2 : (ct >= 13 ? 4 : 8); }
constexpr int gb_tpi(int mode, int ct) { return 512 / (gb_rw(mode, ct) * 16); }
constexpr int gb_exld(int mode, int ct) { return 256 + 32 / gb_tpi(mode, ct); }

// LDS of the back-projection: the staging area (CTINST column tiles, x2 for
// the split modes, or the epilogue's accumulator tiles), the block's ex
// table, (Kx+1) rows of 128 ions x (re, im) padded, and the column scales
template <int MODE>
constexpr size_t g_back_sa_bytes(int CTINST)
{
    return (((size_t)CTINST * 16 * (GK + GLDS_PAD) * (MODE == KM_F64 ? 8 : 4) >
                     (size_t)gb_rw(MODE, CTINST) * 16 * (16 * CTINST + 1) * (MODE == KM_F64 ? 8 : 4)
                 ? (size_t)CTINST * 16 * (GK + GLDS_PAD) * (MODE == KM_F64 ? 8 : 4)
                 : (size_t)gb_rw(MODE, CTINST) * 16 * (16 * CTINST + 1) * (MODE == KM_F64 ? 8 : 4)) +
            15) / 16 * 16;
}

template <int MODE>
inline size_t g_back_lds_bytes(int CTINST, int KX1)
{
    return g_back_sa_bytes<MODE>(CTINST) + (size_t)KX1 * gb_exld(MODE, CTINST) * (MODE == KM_F64 ? 8 : 4) +
           sizeof(double) * GB_KX1_MAX * 6;
}

// back-projection: block = 8 waves x 16 ions, all CT column tiles, one K
// range (grid.y); epilogue: x ex[nx] and the moments per ion.  Grid
// rows from t2_y0 on run back_t2_rows instead (vector-memory work beside
// the matrix-core blocks of the same launch).
template <int MODE, int CTMAX, typename TT>
__global__ void __launch_bounds__(512, ((MODE == KM_F64 || CTMAX >= 13) ? 2 : 4)) k_gemm_back(GBackArgs a)
{
    if (a.active && !*a.active) return;
    if ((int)blockIdx.y >= a.t2_y0) {
        back_t2_rows(a);
        return;
    }
    using T = typename GT<MODE>::T;
    constexpr int LD = GK + GLDS_PAD;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *Bs = (T *)smem_raw;
    T *Bsl = Bs + CTMAX * 16 * LD;   // low parts of the split modes
    // the ex values of this block's ions, [nx][128][re, im], and the
    // column scales [(nx, m)] (1 unless the columns were scaled)
    constexpr int EXLD = gb_exld(MODE, CTMAX);
    TT *exs = (TT *)(smem_raw + g_back_sa_bytes<MODE>(CTMAX));
    const int KX1 = a.NC / (2 * a.CH);
    double *cscs = (double *)(exs + (size_t)KX1 * EXLD);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, cl = lane & 15;
    const int jrow0 = blockIdx.x * 128 + wave * 16;
    const int k0 = blockIdx.y * a.kchunk;
    const int k1 = min(k0 + a.kchunk, a.ldk);
    using acc_t = typename std::conditional<MODE == KM_F64, mf_d4, mf_f4>::type;
    acc_t acc[CTMAX];
#pragma unroll
    for (int c = 0; c < CTMAX; c++) acc[c] = acc_t{0, 0, 0, 0};
    const size_t arow = (size_t)(jrow0 + cl) * a.ldk;
    constexpr bool SPLIT = g_is_split<MODE>();
    constexpr int nslot = CTMAX * 16 * (GK / 8);
    static_assert(nslot <= 512 * GB_SLOTS, "B staging slots");
    // staged with all loads issued first (clamped addresses, no branch
    // around a load), then written to LDS
    constexpr int EXU = (GB_KX1_MAX * 256 + 511) / 512;
    TT xv[EXU];
    const int jb = blockIdx.x * 128, nex = KX1 * 256, ncs = KX1 * a.CH;
#pragma unroll
    for (int u = 0; u < EXU; u++) {
        const int e0 = threadIdx.x + 512 * u, e = e0 < nex ? e0 : 0;
        const int nx = e >> 8, q = e & 255, j0 = jb + (q >> 1), j = j0 < a.nsel ? j0 : 0;
        xv[u] = ((const TT *)a.tx)[2 * ((size_t)nx * a.nsel + j) + (q & 1)];
    }
    const int ics = (int)threadIdx.x < ncs ? (int)threadIdx.x : ncs - 1;
    const double csv = a.csc ? a.csc[ics] : 1.0;
    // two register sets alternate over an even number of steps; loads of
    // step k+2 are issued unconditionally before step k computes (see
    // k_gemm_fwd).  The FP64 form with 13 or more column tiles keeps one set
    // (loads of step k+1 in flight during step k): two do not fit the
    // 256 registers a wave of a 512-thread block can have.
    constexpr bool TWO = !(MODE == KM_F64 && CTMAX >= 13);
    constexpr int NQ = (nslot + 511) / 512;
    // the A runs: loaded (FP64) or formed from the table values of the
    // lane's ion and 4 points (the other modes; ions past nsel repeat the
    // last, points past P the last row, against zero WM columns)
    constexpr bool GEN = MODE != KM_F64;
    struct SetS {
        GRun<MODE> ah, al, rb[NQ], rbl[NQ];
    };
    struct SetG {
        float2 y[4], z[4];
        GRun<MODE> rb[NQ], rbl[NQ];
    };
    using Set = typename std::conditional<GEN, SetG, SetS>::type;
    // register sets in flight: loads of step k+NS are issued before step k
    // computes
    constexpr int NS = TWO ? 2 : 1;
    Set S[NS];
    const int jc = jrow0 + cl < a.nsel ? jrow0 + cl : a.nsel - 1;
    auto load_step = [&](Set &S, int kb) {
        if constexpr (GEN) {
            const int pb = (kb + 8 * g) >> 1;
            int iy = pb / a.NZ, iz = pb - iy * a.NZ;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int yy = iy < a.NY ? iy : a.NY - 1;
                S.y[i] = *(const float2 *)(a.gty + 2 * ((size_t)yy * a.nsel + jc));
                S.z[i] = *(const float2 *)(a.gtz + 2 * ((size_t)iz * a.nsel + jc));
                iz++;
                iy = iz == a.NZ ? iy + 1 : iy;
                iz = iz == a.NZ ? 0 : iz;
            }
        } else {
            S.ah = g_load<MODE>(a.yzt, arow + kb + 8 * g);
            if constexpr (SPLIT) S.al = g_load<MODE>(a.yzt_lo, arow + kb + 8 * g);
        }
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int e0 = threadIdx.x + 512 * q, e = e0 < nslot ? e0 : nslot - 1;
            const int r = e / (GK / 8), q8 = (e % (GK / 8)) * 8;
            S.rb[q] = g_load<MODE>(a.wm, (size_t)r * a.ldk + kb + q8);
            if constexpr (SPLIT) S.rbl[q] = g_load<MODE>(a.wm_lo, (size_t)r * a.ldk + kb + q8);
        }
    };
    auto step = [&](Set &S, int kb, int dist) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int e = threadIdx.x + 512 * q;
            if (e < nslot) {
                const int r = e / (GK / 8), q8 = (e % (GK / 8)) * 8;
                *(decltype(S.rb[q].v) *)(Bs + r * LD + q8) = S.rb[q].v;
                if constexpr (SPLIT) *(decltype(S.rbl[q].v) *)(Bsl + r * LD + q8) = S.rbl[q].v;
            }
        }
        GRun<MODE> ch, cl_;
        if constexpr (GEN) g_gen_run<MODE>(S.y, S.z, ch, cl_);
        else { ch = S.ah; cl_ = S.al; }
        __syncthreads();
        load_step(S, kb + dist * GK < k1 ? kb + dist * GK : kb);
        if constexpr (SPLIT) {
            using V = gvec_t<MODE>;
#pragma unroll
            for (int ct = 0; ct < CTMAX; ct++) {
                const V bh = *(const V *)(Bs + (ct * 16 + cl) * LD + 8 * g);
                const V bl = *(const V *)(Bsl + (ct * 16 + cl) * LD + 8 * g);
                acc[ct] = g_mma_split<MODE>(ch.v, bh, acc[ct]);
                acc[ct] = g_mma_split<MODE>(ch.v, bl, acc[ct]);
                acc[ct] = g_mma_split<MODE>(cl_.v, bh, acc[ct]);
            }
        } else {
#pragma unroll
            for (int s = 0; s < 8; s++)
#pragma unroll
                for (int ct = 0; ct < CTMAX; ct++)
                    acc[ct] = mf_mma<MODE>(ch.v[s], Bs[(ct * 16 + cl) * LD + 8 * g + s], acc[ct]);
        }
    };
    if (k0 < k1) {
        load_step(S[0], k0);
        if constexpr (NS > 1) load_step(S[1], k0 + GK < k1 ? k0 + GK : k0);
    }
    // the staged tables into LDS (their loads were issued before the
    // first steps' loads); read in the epilogue only
#pragma unroll
    for (int u = 0; u < EXU; u++) {
        const int e = threadIdx.x + 512 * u;
        if (e < nex) exs[(e >> 8) * EXLD + (e & 255)] = (jb + ((e & 255) >> 1)) < a.nsel ? xv[u] : (TT)0;
    }
    if ((int)threadIdx.x < ncs) cscs[threadIdx.x] = csv;
    for (int kb = k0; kb < k1; kb += NS * GK) {
        step(S[0], kb, NS);
        if constexpr (NS > 1) step(S[1], kb + GK, NS);
    }
    __syncthreads();   // the staging area is free; the ex table is complete (also for an empty K range)
    // Epilogue, RW waves at a time (all 8 where the LDS allows): their
    // accumulator tiles (16 ions x all columns each) go to the staging area,
    // then TPI threads per ion take every TPI-th nx and form Z = ex[nx] T for
    // each moment m (a uniform, unrolled loop: no lane-dependent branch), and
    // the TPI lanes of an ion are summed.  Outputs: field H (x nx, 1, 1 for
    // m = 0, 1, 2) and, with the gradient, M (x nx^2, nx, nx, 1, 1, 1 for
    // m = 0..5).
    using et = typename std::conditional<MODE == KM_F64, double, float>::type;
    constexpr int TLD = 16 * CTMAX + 1;
    constexpr int RW = gb_rw(MODE, CTMAX), TPI = gb_tpi(MODE, CTMAX);
    et *tile = (et *)smem_raw;
    const int nch = (a.CH == 6) ? 9 : 3;
    double *out = a.part + (size_t)blockIdx.y * nch * a.nsel;
    const int er = threadIdx.x / TPI, es = threadIdx.x % TPI;
    for (int round = 0; round < 8 / RW; round++) {
        if (wave / RW == round) {
            et *tw = tile + (wave % RW) * 16 * TLD;
#pragma unroll
            for (int ct = 0; ct < CTMAX; ct++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int r = (MODE == KM_F64) ? g + 4 * i : 4 * g + i;
                    tw[r * TLD + ct * 16 + cl] = acc[ct][i];
                }
        }
        __syncthreads();
        const int jl = round * RW * 16 + er;         // ion within the block
        double o[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        // every nx plane of this thread (at most GB_KX1_MAX / TPI, rounded
        // up), unrolled so that the LDS reads of all planes are issued first
        constexpr int NXI = (GB_KX1_MAX + TPI - 1) / TPI;
        if (a.CH == 3) {
            // field only: Im z of the three moments
#pragma unroll
            for (int it = 0; it < NXI; it++) {
                const int nx = es + it * TPI;
                if (nx < KX1) {
                    const double xr = (double)exs[nx * EXLD + 2 * jl], xi = (double)exs[nx * EXLD + 2 * jl + 1];
                    const et *rowp = tile + er * TLD + nx * 6;
                    const double *cs = cscs + nx * 3;
                    double zi[3];
#pragma unroll
                    for (int m = 0; m < 3; m++)
                        zi[m] = (xr * (double)rowp[2 * m + 1] + xi * (double)rowp[2 * m]) * cs[m];
                    o[0] += (double)nx * zi[0];
                    o[1] += zi[1];
                    o[2] += zi[2];
                }
            }
        } else {
#pragma unroll
            for (int it = 0; it < NXI; it++) {
                const int nx = es + it * TPI;
                if (nx < KX1) {
                    const double xr = (double)exs[nx * EXLD + 2 * jl], xi = (double)exs[nx * EXLD + 2 * jl + 1];
                    const double fnx = (double)nx;
                    const et *rowp = tile + er * TLD + nx * 12;
                    const double *cs = cscs + nx * 6;
                    auto z = [&](int m, double &zr, double &zi) {
                        const double tre = (double)rowp[2 * m] * cs[m], tim = (double)rowp[2 * m + 1] * cs[m];
                        zr = xr * tre - xi * tim;
                        zi = xr * tim + xi * tre;
                    };
                    double zr, zi;
                    z(0, zr, zi); o[0] += fnx * zi; o[3] += fnx * fnx * zr;
                    z(1, zr, zi); o[1] += zi; o[4] += fnx * zr;
                    z(2, zr, zi); o[2] += zi; o[5] += fnx * zr;
                    z(3, zr, zi); o[6] += zr;
                    z(4, zr, zi); o[7] += zr;
                    z(5, zr, zi); o[8] += zr;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 9; c++) {
            if (c >= nch) continue;
            double v = o[c];
#pragma unroll
            for (int w = 1; w < TPI; w <<= 1) v += __shfl_xor(v, w, 64);
            o[c] = v;
        }
        const int j = blockIdx.x * 128 + jl;
        if (es == 0 && j < a.nsel) {
#pragma unroll
            for (int c = 0; c < 9; c++)
                if (c < nch) out[(size_t)c * a.nsel + j] = o[c];
        }
        __syncthreads();   // the tiles are rewritten in the next round
    }
}
