// batch_kernels.h -- kernels of the batched engine: many configurations of
// uneven size in concatenated SoA arrays, each with its own cell, Ewald
// alpha, k box and cutoffs (BConfD).  Real-space forces reuse k_short and
// k_ewald_dir with BATCH = true; this file holds the per-configuration
// neighbour list, the k-space sums over per-configuration k lists, the
// segmented reductions, the SCF steps gated per configuration, and the
// adjoint kernels of the force-matching loss.
#pragma once
#include "dev.h"

// one half-space k-vector of a configuration (kspace.c:14-102)
struct KVec {
    double kx, ky, kz, Ak, kcoe;
    int nx, ny, nz, pad;
};

// a (configuration, tile) work item of the persistent k-space kernels
struct WItem {
    int conf, tile;
};

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ bool conf_active(const int *act, int c) { return !act || act[c]; }

// Work queue of a persistent grid: thread 0 takes the next item; the last
// block to run dry resets the counters for the next launch on the stream.
__device__ __forceinline__ int next_item(int *queue, int nitems, int *sh)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        int it = atomicAdd(&queue[0], 1);
        if (it >= nitems) {
            __threadfence();
            int done = atomicAdd(&queue[1], 1);
            if (done == (int)gridDim.x - 1) {
                queue[0] = 0;
                queue[1] = 0;
                __threadfence();
            }
        }
        *sh = it;
    }
    __syncthreads();
    return *sh;
}

// kardir cell.c:125-143 with each ion's own cell
__global__ void kb_frac(int n, const BConfD *__restrict__ conf, const int *__restrict__ cfg,
                        const double *__restrict__ r, double *__restrict__ s)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const CellD &c = conf[cfg[i]].c;
    double x = r[i], y = r[n + i], z = r[2 * n + i];
    s[i] = x * c.B[0][0] + y * c.B[0][1] + z * c.B[0][2];
    s[n + i] = x * c.B[1][0] + y * c.B[1][1] + z * c.B[1][2];
    s[2 * n + i] = x * c.B[2][0] + y * c.B[2][1] + z * c.B[2][2];
}

// out[c][col] = sum over the ions (kspace = 0) or k-vectors (kspace = 1) of
// configuration c of in[col*ld + .]; one block per (c, col), fixed order.
__global__ void __launch_bounds__(256) kb_segsum(const double *__restrict__ in, size_t ld, int ncol,
                                                 const BConfD *__restrict__ conf, int kspace,
                                                 double *__restrict__ out)
{
    __shared__ double sh[256];
    const int c = blockIdx.x, col = blockIdx.y;
    const int b = kspace ? conf[c].koff : conf[c].off;
    const int m = kspace ? conf[c].nk : conf[c].n;
    const double *p = in + (size_t)col * ld + b;
    double s = 0.0;
    for (int i = threadIdx.x; i < m; i += 256) s += p[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)c * ncol + col] = sh[0];
}

// out[t][col] = sum over ions of type t of in[col*ld + i]
__global__ void __launch_bounds__(256) kb_typesum(const double *__restrict__ in, size_t ld, int ncol, int n,
                                                  const int *__restrict__ type, double *__restrict__ out)
{
    __shared__ double sh[256];
    const int t = blockIdx.x, col = blockIdx.y;
    const double *p = in + (size_t)col * ld;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256)
        if (type[i] == t) s += p[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)t * ncol + col] = sh[0];
}

// full neighbour list of each ion within its configuration (one wavefront
// per ion, ordered by j): entry j | NBR_SR (d^2 <= rc_s^2) | NBR_LR
// (d^2 <= rc_l^2); ctl[0] = largest count beyond cap (overflow)
__global__ void __launch_bounds__(256) kb_list(int n, const BConfD *__restrict__ conf, const int *__restrict__ cfg,
                                               const double *__restrict__ s, int cap, unsigned *__restrict__ nbr,
                                               int *__restrict__ cnt, int *ctl)
{
    int lane = threadIdx.x & 63;
    int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const BConfD &K = conf[cfg[i]];
    const double six = s[i], siy = s[n + i], siz = s[2 * n + i];
    int count = 0;
    for (int j0 = K.off; j0 < K.off + K.n; j0 += 64) {
        int j = j0 + lane;
        unsigned e = 0;
        if (j < K.off + K.n && j != i) {
            double x, y, z;
            min_image(K.c, six - s[j], siy - s[n + j], siz - s[2 * n + j], x, y, z);
            double d2 = x * x + y * y + z * z;
            if (d2 <= K.rc2s) e |= NBR_SR;
            if (d2 <= K.rc2l) e |= NBR_LR;
            if (e) e |= (unsigned)j;
        }
        unsigned long long m = __ballot(e != 0);
        int pos = count + __popcll(m & ((1ull << lane) - 1ull));
        if (e != 0 && pos < cap) nbr[(size_t)i * cap + pos] = e;
        count += __popcll(m);
    }
    if (lane == 0) {
        cnt[i] = count < cap ? count : cap;
        if (count > cap) atomicMax(&ctl[0], count);
    }
}

// --------------------------------------------------------------- k space
// e^{i n theta_b} tables per ion, AoS with a fixed stride of TS complex
// numbers: x at [0, Kx], y centred at by (n = -Ky..Ky), z centred at bz;
// theta_b = 2 pi t_b, t_b = sum_a B[a][b]... as kardir, wrapped to [-1/2, 1/2].
// tabT (when given) receives the same values column by column, complex
// entry m of ion i at [m][i] (coalesced over consecutive ions).
template <typename real>
__global__ void kb_eikr(int n, const BConfD *__restrict__ conf, const int *__restrict__ cfg,
                        const double *__restrict__ r, int TS, int by, int bz, real *__restrict__ tab,
                        real *__restrict__ tabT = nullptr)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const BConfD &K = conf[cfg[i]];
    double x = r[i], y = r[n + i], z = r[2 * n + i];
    double t[3];
    for (int b = 0; b < 3; b++) {
        double v = x * K.c.B[0][b] + y * K.c.B[1][b] + z * K.c.B[2][b];
        t[b] = v - rint(v);
    }
    real *T = tab + 2 * (size_t)i * TS;
    double th = MDFF_TPI * t[0];
    for (int m = 0; m <= K.Kx; m++) {
        double sn, cs;
        sincos(th * (double)m, &sn, &cs);
        T[2 * m] = (real)cs;
        T[2 * m + 1] = (real)sn;
    }
    th = MDFF_TPI * t[1];
    for (int m = -K.Ky; m <= K.Ky; m++) {
        double sn, cs;
        sincos(th * (double)m, &sn, &cs);
        T[2 * (by + m)] = (real)cs;
        T[2 * (by + m) + 1] = (real)sn;
    }
    th = MDFF_TPI * t[2];
    for (int m = -K.Kz; m <= K.Kz; m++) {
        double sn, cs;
        sincos(th * (double)m, &sn, &cs);
        T[2 * (bz + m)] = (real)cs;
        T[2 * (bz + m) + 1] = (real)sn;
    }
    if (tabT) {
        // the columns this configuration uses (x 0..Kx, y and z -K..K)
        typedef real r2 __attribute__((ext_vector_type(2)));
        auto col = [&](int m) { *(r2 *)(tabT + 2 * ((size_t)m * n + i)) = *(const r2 *)(T + 2 * m); };
        for (int m = 0; m <= K.Kx; m++) col(m);
        for (int m = -K.Ky; m <= K.Ky; m++) col(by + m);
        for (int m = -K.Kz; m <= K.Kz; m++) col(bz + m);
    }
}

// weights of the charge density rho(k) = sum_j w_j(k) e^{ik.r_j}
enum {
    RM_Q = 0,    // w = q                                   (charge_density_q kspace.c:189-207)
    RM_MU = 1,   // w = i k.mu                              (charge_density_mu :209-230)
    RM_QMU = 2,  // w = q + i k.mu                          (charge_density_qmu :232-256)
    RM_ADJ = 3   // w = i k.(q g) - (k.mu)(k.g): the change of rho when ion j moves by g_j
};

struct RhoArgs {
    const BConfD *conf;
    const KVec *kv;
    const void *tab;
    int TS, by, bz;
    int n;                     // ions in the batch (SoA stride)
    const double *q, *mu, *g;  // weights: q[n], mu[3n], g[3n] (RM_ADJ)
    const WItem *items;
    int nitems;
    int *queue;
    double *W;                 // [nk][2] = Ak conj(rho)
    double *ek;                // [7][nk]: Ak |rho|^2 and the 6 stress terms
    size_t nktot;
    int do_stress;
    const int *act;            // per configuration (may be null)
};

// rho(k), str = Ak |rho|^2 and the stress terms of multipole_ES_rec
// (ewald.c:949-1006) for one (configuration, KPB k-vectors) item per block
// pass; the 256/KPB thread slices split the configuration's ions and are
// summed in a fixed order.
template <typename real, int MODE, int KPB>
__global__ void __launch_bounds__(256) kb_rho(RhoArgs a)
{
    constexpr int S = 256 / KPB;
    __shared__ double shr[S][KPB], shi[S][KPB];
    __shared__ int s_item;
    const int kl = threadIdx.x % KPB, sl = threadIdx.x / KPB;
    const real *tab = (const real *)a.tab;
    const int n = a.n;
    for (;;) {
        int it = next_item(a.queue, a.nitems, &s_item);
        if (it >= a.nitems) break;
        const WItem w = a.items[it];
        const BConfD &K = a.conf[w.conf];
        if (!conf_active(a.act, w.conf)) continue;
        const int kidx = w.tile * KPB + kl;
        const bool valid = kidx < K.nk;
        double rr = 0.0, ri = 0.0;
        if (valid) {
            const KVec kv = a.kv[K.koff + kidx];
            const int ox = 2 * kv.nx, oy = 2 * (a.by + kv.ny), oz = 2 * (a.bz + kv.nz);
            for (int j = K.off + sl; j < K.off + K.n; j += S) {
                const real *T = tab + 2 * (size_t)j * a.TS;
                double xr = T[ox], xi = T[ox + 1], yr = T[oy], yi = T[oy + 1], zr = T[oz], zi = T[oz + 1];
                double pr = xr * yr - xi * yi, pi = xr * yi + xi * yr;
                double er = pr * zr - pi * zi, ei = pr * zi + pi * zr;
                double wr = 0.0, wi = 0.0;
                if (MODE == RM_Q || MODE == RM_QMU) wr = a.q[j];
                if (MODE == RM_MU || MODE == RM_QMU)
                    wi = kv.kx * a.mu[j] + kv.ky * a.mu[n + j] + kv.kz * a.mu[2 * n + j];
                if (MODE == RM_ADJ) {
                    double kg = kv.kx * a.g[j] + kv.ky * a.g[n + j] + kv.kz * a.g[2 * n + j];
                    double km = kv.kx * a.mu[j] + kv.ky * a.mu[n + j] + kv.kz * a.mu[2 * n + j];
                    wr = -km * kg;
                    wi = a.q[j] * kg;
                }
                rr += wr * er - wi * ei;
                ri += wr * ei + wi * er;
            }
        }
        shr[sl][kl] = rr;
        shi[sl][kl] = ri;
        __syncthreads();
        if (sl == 0 && valid) {
            double RR = 0.0, RI = 0.0;
            for (int t = 0; t < S; t++) { RR += shr[t][kl]; RI += shi[t][kl]; }
            const KVec kv = a.kv[K.koff + kidx];
            const size_t k = (size_t)K.koff + kidx;
            a.W[2 * k] = kv.Ak * RR;
            a.W[2 * k + 1] = -kv.Ak * RI;
            if (a.ek) {
                double str = (RR * RR + RI * RI) * kv.Ak;
                const size_t nk = a.nktot;
                a.ek[k] = str;
                double kc = kv.kcoe;
                double v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0, v6 = 0;
                if (a.do_stress) {
                    v1 = (1.0 - kc * kv.kx * kv.kx) * str; v2 = -kc * kv.kx * kv.ky * str;
                    v3 = -kc * kv.kx * kv.kz * str; v4 = (1.0 - kc * kv.ky * kv.ky) * str;
                    v5 = -kc * kv.ky * kv.kz * str; v6 = (1.0 - kc * kv.kz * kv.kz) * str;
                }
                a.ek[nk + k] = v1; a.ek[2 * nk + k] = v2; a.ek[3 * nk + k] = v3;
                a.ek[4 * nk + k] = v4; a.ek[5 * nk + k] = v5; a.ek[6 * nk + k] = v6;
            }
        }
    }
}

struct BackBArgs {
    const BConfD *conf;
    const KVec *kv;
    const void *tab;       // tables column by column, [TS][n] complex (kb_eikr's tabT)
    int TS, by, bz;
    int n;
    const double *W;       // [nk][2]
    const WItem *items;
    int nitems;
    int *queue;
    const double *q, *mu;  // force weights
    double *ef;            // [3][n]  8 pi/V sum_k k Im(W e^{ik.r})
    double *efg;           // [6][n]  8 pi/V sum_k k k Re(W e^{ik.r}) (xx xy xz yy yz zz), NCH = 9
    double *f;             // [3][n]  8 pi/V (q E + G mu) when non-null (NCH = 9 for dipoles)
    const int *act;
};

// field, field gradient and force of multipole_ES_rec (ewald.c:957-1034) for
// one (configuration, IPB ions) item per block pass; the 256/IPB slices split
// the configuration's k-vectors and are summed in a fixed order.
template <typename real, int NCH, int IPB>
__global__ void __launch_bounds__(256) kb_back(BackBArgs a)
{
    constexpr int S = 256 / IPB;
    __shared__ double sh[S][NCH][IPB];
    __shared__ int s_item;
    const int il = threadIdx.x % IPB, sl = threadIdx.x / IPB;
    const real *tab = (const real *)a.tab;
    const int n = a.n;
    for (;;) {
        int it = next_item(a.queue, a.nitems, &s_item);
        if (it >= a.nitems) break;
        const WItem w = a.items[it];
        const BConfD &K = a.conf[w.conf];
        if (!conf_active(a.act, w.conf)) continue;
        const int li = w.tile * IPB + il;
        const bool valid = li < K.n;
        const int i = K.off + li;
        double acc[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++) acc[c] = 0.0;
        if (valid) {
            typedef real r2 __attribute__((ext_vector_type(2)));
            const r2 *T = (const r2 *)tab + i;
            for (int k = sl; k < K.nk; k += S) {
                const KVec kv = a.kv[K.koff + k];
                const double Wr = a.W[2 * ((size_t)K.koff + k)], Wi = a.W[2 * ((size_t)K.koff + k) + 1];
                const r2 X = T[(size_t)kv.nx * n], Y = T[(size_t)(a.by + kv.ny) * n], Z = T[(size_t)(a.bz + kv.nz) * n];
                double xr = X[0], xi = X[1], yr = Y[0], yi = Y[1], zr = Z[0], zi = Z[1];
                double pr = xr * yr - xi * yi, pi = xr * yi + xi * yr;
                double er = pr * zr - pi * zi, ei = pr * zi + pi * zr;
                double zre = Wr * er - Wi * ei, zim = Wr * ei + Wi * er;
                acc[0] += kv.kx * zim;
                acc[1] += kv.ky * zim;
                acc[2] += kv.kz * zim;
                if (NCH == 9) {
                    acc[3] += kv.kx * kv.kx * zre;
                    acc[4] += kv.kx * kv.ky * zre;
                    acc[5] += kv.kx * kv.kz * zre;
                    acc[6] += kv.ky * kv.ky * zre;
                    acc[7] += kv.ky * kv.kz * zre;
                    acc[8] += kv.kz * kv.kz * zre;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < NCH; c++) sh[sl][c][il] = acc[c];
        __syncthreads();
        if (sl == 0 && valid) {
            double v[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                double s = 0.0;
                for (int t = 0; t < S; t++) s += sh[t][c][il];
                v[c] = s * K.tfpiV;
            }
            if (a.ef) { a.ef[i] = v[0]; a.ef[n + i] = v[1]; a.ef[2 * n + i] = v[2]; }
            if (NCH == 9 && a.efg) {
#pragma unroll
                for (int c = 0; c < 6; c++) a.efg[(size_t)c * n + i] = v[3 + c];
            }
            if (a.f) {
                double qi = a.q[i];
                double fx = qi * v[0], fy = qi * v[1], fz = qi * v[2];
                if (NCH == 9) {
                    double m0 = a.mu[i], m1 = a.mu[n + i], m2 = a.mu[2 * n + i];
                    fx += v[3] * m0 + v[4] * m1 + v[5] * m2;
                    fy += v[4] * m0 + v[6] * m1 + v[7] * m2;
                    fz += v[5] * m0 + v[7] * m1 + v[8] * m2;
                }
                a.f[i] = fx; a.f[n + i] = fy; a.f[2 * n + i] = fz;
            }
        }
    }
}

// ------------------------------------------- matrix-core k-space sums
// The same sums as kb_rho / kb_back, recast per configuration as dense
// GEMMs over its full k box (nx = 0..Kx, ny = -Ky..Ky, nz = -Kz..Kz; box
// point p = (ny + Ky) NZ + (nz + Kz), NZ = 2 Kz + 1) on the separable
// tables e^{ik.r} = ex[nx] ey[ny] ez[nz] (ewald.c:876-1051 semantics):
//   forward  S_c(nx, p) = sum_j w_cj ex_j[nx] (ey ez)_j(p): rows (nx, c,
//            re/im), columns p, K = (ion, re/im);
//   back     T_m(j, nx) = sum_p (ey ez)_j(p) W(nx, p) mom_m(p): rows ions,
//            columns (nx, m, re/im), K = (point, re/im), moments
//            mom = {1, ny, nz, ny^2, ny nz, nz^2}; then ex_j[nx] and the nx
//            factors per ion.
// Complex products are embedded as real GEMMs: for a = ar + i ai times
// b = br + i bi over K = (re, im), the row (re) takes (ar, -ai) and the row
// (im) takes (ai, ar).  The operand that depends on the weights (forward:
// w ex rows; back: moment-weighted W columns) is written once per call in
// the embedded, K-major layout the lanes read (kb_mf_aw, kb_mf_wm); the
// other ((ey ez)_j(p)) is formed in registers.  One wave computes one
// item: a configuration's column group against all its row tiles (forward)
// or R 16-ion row tiles against all its column tiles (back), over the
// whole K range: no partial sums.  A launch covers the configurations of
// one tile count (compile-time MT / CT).  v_mfma_f64_16x16x4f64 (FP64 form,
// and the mixed form's SCF fields) and v_mfma_f32_16x16x4f32 (the mixed
// form's static field and force call): A[l % 16][k], B[k][l % 16] with
// k = l / 16.

template <typename real>
struct BMF;
template <>
struct BMF<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row of accumulator element e held by lane l (whose column is l % 16)
    static __device__ __forceinline__ int row(int l, int e) { return (l >> 4) + 4 * e; }
};
template <>
struct BMF<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int e) { return 4 * (l >> 4) + e; }
};

#define BMF_MT 8    // row tiles of the forward sum: (Kx + 1) C 2 <= 16 BMF_MT
#define BMF_CT 12   // column tiles of the back-projection: (Kx + 1) 6 2 <= 16 BMF_CT

// tile counts of a configuration, rounded up to even (the launch groups)
__host__ __device__ __forceinline__ int bmf_mt(int Kx, int C) { return ((Kx + 1) * C * 2 + 31) / 32 * 2; }
__host__ __device__ __forceinline__ int bmf_ct(int Kx, int CH) { return ((Kx + 1) * CH * 2 + 31) / 32 * 2; }

// forward A operand: AW[awoff + kk Rp + row] (kk = 2 j_local + re/im, Rp =
// 16 bmf_mt(Kx, C)), rows past (Kx + 1) C 2 zero; grid (chunks,
// configurations)
template <typename TT, typename AT>
__global__ void __launch_bounds__(256) kb_mf_aw(const BConfD *__restrict__ conf, const void *__restrict__ tabT,
                                                int n, int C, const double *__restrict__ q,
                                                const double *__restrict__ mu, AT *__restrict__ aw,
                                                const int *__restrict__ act)
{
    const int c = blockIdx.y;
    if (!conf_active(act, c)) return;
    const BConfD &K = conf[c];
    typedef TT t2 __attribute__((ext_vector_type(2)));
    const t2 *T = (const t2 *)tabT;
    const int R = (K.Kx + 1) * C * 2, Rp = 16 * bmf_mt(K.Kx, C);
    const int tot = 2 * K.n * Rp;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += gridDim.x * blockDim.x) {
        const int kk = e / Rp, row = e - kk * Rp;
        AT v = (AT)0;
        if (row < R) {
            const int j = K.off + (kk >> 1), nx = row / (2 * C), ch = (row >> 1) % C;
            const int wc = C == 4 ? ch : (C == 3 ? ch + 1 : 0);   // 0 q, 1..3 mu
            const double w = wc == 0 ? q[j] : mu[(size_t)(wc - 1) * n + j];
            const t2 X = T[(size_t)nx * n + j];
            const double ar = w * (double)X.x, ai = w * (double)X.y;
            const bool kim = kk & 1, rim = row & 1;
            v = (AT)(rim ? (kim ? ar : ai) : (kim ? -ai : ar));
        }
        aw[(size_t)K.awoff + e] = v;
    }
}

// back B operand: WM[wmoff + kk NCp + col] (kk = 2 p + re/im over the box
// points rounded up to even, NCp = 16 bmf_ct(Kx, CH)), the moment-weighted W
// of the half-space k-vectors (zero elsewhere and past (Kx + 1) CH 2)
template <typename WT>
__global__ void __launch_bounds__(256) kb_mf_wm(const BConfD *__restrict__ conf, const int *__restrict__ boxk,
                                                const double *__restrict__ W, int CH, WT *__restrict__ wm,
                                                const int *__restrict__ act)
{
    const int c = blockIdx.y;
    if (!conf_active(act, c)) return;
    const BConfD &K = conf[c];
    const int NZ = 2 * K.Kz + 1, P = (2 * K.Ky + 1) * NZ, Pp = (P + 1) & ~1;
    const int NC = (K.Kx + 1) * CH * 2, NCp = 16 * bmf_ct(K.Kx, CH);
    const int tot = 2 * Pp * NCp;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += gridDim.x * blockDim.x) {
        const int kk = e / NCp, col = e - kk * NCp, pp = kk >> 1;
        WT v = (WT)0;
        if (col < NC && pp < P) {
            const int nx = col / (2 * CH), m = (col >> 1) % CH;
            const int k = boxk[K.boxoff + nx * P + pp];
            if (k >= 0) {
                const int iy = pp / NZ;
                const double ny = (double)(iy - K.Ky), nz = (double)(pp - iy * NZ - K.Kz);
                const double mom = m == 0 ? 1.0 : (m == 1 ? ny : (m == 2 ? nz : (m == 3 ? ny * ny : (m == 4 ? ny * nz : nz * nz))));
                const double wr = mom * W[2 * ((size_t)K.koff + k)], wi = mom * W[2 * ((size_t)K.koff + k) + 1];
                const bool kim = kk & 1, cim = col & 1;
                v = (WT)(cim ? (kim ? wr : wi) : (kim ? -wi : wr));
            }
        }
        wm[(size_t)K.wmoff + e] = v;
    }
}

struct MFwdArgs {
    const BConfD *conf;
    const void *tabT;          // [TS][n] complex (TT), column m of ion i at m n + i
    const void *aw;            // A operands (kb_mf_aw)
    int by, bz, n, C;          // C: 1 (q), 3 (mu), 4 (q, mu)
    const int *boxk;           // [box point (nx, p)] -> k index within the configuration, or -1
    const WItem *items;        // (configuration, column group of CW tiles), this launch's range
    int nitems;
    double *Sk;                // [nktot][C][re, im]
    const int *act;
};

// TT: table type; AT: A operand type; real: arithmetic of the matrix
// instructions; MT row tiles (every item of the launch has bmf_mt = MT);
// CW column tiles per item.  Operands of step s + 1 load while step s
// computes.
template <typename TT, typename AT, typename real, int MT, int CW>
__global__ void __launch_bounds__(256) kb_mf_fwd(MFwdArgs a)
{
    using M = BMF<real>;
    using v4 = typename M::v4;
    typedef TT t2 __attribute__((ext_vector_type(2)));
    const int lane = threadIdx.x & 63;
    const int it = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (it >= a.nitems) return;
    const WItem w = a.items[it];
    if (!conf_active(a.act, w.conf)) return;
    const BConfD &K = a.conf[w.conf];
    const int g = lane >> 4, cl = lane & 15;
    const int C = a.C, NZ = 2 * K.Kz + 1, P = (2 * K.Ky + 1) * NZ;
    const int R = (K.Kx + 1) * C * 2;
    constexpr int Rp = 16 * MT;
    const int n = a.n;
    const t2 *T = (const t2 *)a.tabT;
    // the lane's A column (row of the sum) in the AW block and its B columns
    // (ny, nz); columns past the box repeat the last one (dropped)
    const AT *A = (const AT *)a.aw + K.awoff + cl;
    int yo[CW], zo[CW];
#pragma unroll
    for (int c = 0; c < CW; c++) {
        const int p = min((w.tile * CW + c) * 16 + cl, P - 1);
        const int iy = p / NZ;
        yo[c] = (a.by - K.Ky + iy) * n;
        zo[c] = (a.bz - K.Kz + (p - iy * NZ)) * n;
    }
    v4 acc[MT][CW];
#pragma unroll
    for (int t = 0; t < MT; t++)
#pragma unroll
        for (int c = 0; c < CW; c++) acc[t][c] = v4{0, 0, 0, 0};
    const bool kim = g & 1;   // the lane's K component: im
    const int ns = (K.n + 1) >> 1;
    struct Set {
        t2 Y[CW], Z[CW];
        AT X[MT];
    };
    // step s: ions 2 s, 2 s + 1 (the last ion repeated past the configuration,
    // against zero B)
    auto load = [&](Set &S, int s) {
        const int jl = min(2 * s + (g >> 1), K.n - 1);
        const int j = K.off + jl;
#pragma unroll
        for (int c = 0; c < CW; c++) {
            S.Y[c] = T[yo[c] + j];
            S.Z[c] = T[zo[c] + j];
        }
        const AT *Ar = A + (size_t)(2 * jl + (g & 1)) * Rp;   // K row 2 jl + re/im of the block
#pragma unroll
        for (int t = 0; t < MT; t++) S.X[t] = Ar[t * 16];
    };
    auto compute = [&](const Set &S, int s) {
        const bool jok = 2 * s + (g >> 1) < K.n;
        real bv[CW];
#pragma unroll
        for (int c = 0; c < CW; c++) {
            const real yr = (real)S.Y[c].x, yi = (real)S.Y[c].y, zr = (real)S.Z[c].x, zi = (real)S.Z[c].y;
            const real re = yr * zr - yi * zi, im = yr * zi + yi * zr;
            bv[c] = jok ? (kim ? im : re) : (real)0;
        }
#pragma unroll
        for (int t = 0; t < MT; t++) {
            const real av = (real)S.X[t];
#pragma unroll
            for (int c = 0; c < CW; c++) acc[t][c] = M::mma(av, bv[c], acc[t][c]);
        }
    };
    Set S0, S1;
    load(S0, 0);
    int s = 0;
    for (; s + 1 < ns; s += 2) {
        load(S1, s + 1);
        compute(S0, s);
        load(S0, min(s + 2, ns - 1));
        compute(S1, s + 1);
    }
    if (s < ns) compute(S0, s);
    // S values of the box points that are half-space k-vectors
    const int *bk = a.boxk + K.boxoff;
#pragma unroll
    for (int t = 0; t < MT; t++) {
#pragma unroll
        for (int c = 0; c < CW; c++) {
            const int p = (w.tile * CW + c) * 16 + cl;
            if (p >= P) continue;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = t * 16 + M::row(lane, e);
                if (r >= R) continue;
                const int k = bk[(r / (2 * C)) * P + p];
                if (k >= 0) a.Sk[((size_t)(K.koff + k) * C + ((r >> 1) % C)) * 2 + (r & 1)] = (double)acc[t][c][e];
            }
        }
    }
}

// rho(k) of every half-space k-vector from the box sums, then W = Ak
// conj(rho), Ak |rho|^2 and the stress terms as kb_rho (C = 1: RM_Q, 3:
// RM_MU, 4: RM_QMU)
__global__ void __launch_bounds__(256) kb_mf_rho(int nktot, const int *__restrict__ kcfg,
                                                 const KVec *__restrict__ kv, const double *__restrict__ Sk, int C,
                                                 double *__restrict__ W, double *__restrict__ ek, int do_stress,
                                                 const int *__restrict__ act)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nktot) return;
    if (!conf_active(act, kcfg[k])) return;
    const KVec v = kv[k];
    const double *s = Sk + (size_t)k * C * 2;
    double RR, RI;
    if (C == 1) {
        RR = s[0];
        RI = s[1];
    } else if (C == 3) {
        RR = -(v.kx * s[1] + v.ky * s[3] + v.kz * s[5]);
        RI = v.kx * s[0] + v.ky * s[2] + v.kz * s[4];
    } else {
        RR = s[0] - (v.kx * s[3] + v.ky * s[5] + v.kz * s[7]);
        RI = s[1] + (v.kx * s[2] + v.ky * s[4] + v.kz * s[6]);
    }
    W[2 * (size_t)k] = v.Ak * RR;
    W[2 * (size_t)k + 1] = -v.Ak * RI;
    if (ek) {
        const size_t nk = (size_t)nktot;
        const double str = (RR * RR + RI * RI) * v.Ak;
        ek[k] = str;
        const double kc = v.kcoe;
        double v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0, v6 = 0;
        if (do_stress) {
            v1 = (1.0 - kc * v.kx * v.kx) * str; v2 = -kc * v.kx * v.ky * str;
            v3 = -kc * v.kx * v.kz * str; v4 = (1.0 - kc * v.ky * v.ky) * str;
            v5 = -kc * v.ky * v.kz * str; v6 = (1.0 - kc * v.kz * v.kz) * str;
        }
        ek[nk + k] = v1; ek[2 * nk + k] = v2; ek[3 * nk + k] = v3;
        ek[4 * nk + k] = v4; ek[5 * nk + k] = v5; ek[6 * nk + k] = v6;
    }
}

struct MBackArgs {
    const BConfD *conf;
    const void *tabT;          // [TS][n] complex (TT)
    const void *wm;            // B operands (kb_mf_wm)
    int by, bz, n, CH;         // CH: 3 moments (field) or 6 (field and gradient)
    const WItem *items;        // (configuration, group of R 16-ion row tiles), this launch's range
    int nitems;
    const double *q, *mu;
    double *ef, *efg, *f;      // as kb_back
    const int *act;
};

// TT: table type; WT: B operand type; real: arithmetic of the matrix
// instructions; CT column tiles (every item of the launch has bmf_ct = CT);
// R row tiles per item.  Operands of step s + 1 load while step s computes.
template <typename TT, typename WT, typename real, int CT, int R>
__global__ void __launch_bounds__(256) kb_mf_back(MBackArgs a)
{
    using M = BMF<real>;
    using v4 = typename M::v4;
    typedef TT t2 __attribute__((ext_vector_type(2)));
    const int lane = threadIdx.x & 63;
    const int it = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (it >= a.nitems) return;
    const WItem w = a.items[it];
    if (!conf_active(a.act, w.conf)) return;
    const BConfD &K = a.conf[w.conf];
    const int g = lane >> 4, cl = lane & 15;
    const int CH = a.CH, NZ = 2 * K.Kz + 1, P = (2 * K.Ky + 1) * NZ;
    const int NC = (K.Kx + 1) * CH * 2;
    constexpr int NCp = 16 * CT;
    const int n = a.n;
    const t2 *T = (const t2 *)a.tabT;
    // the lane's B column in the WM block (K row g of each step)
    const WT *B = (const WT *)a.wm + K.wmoff + g * NCp + cl;
    // A rows: the lane's ion in each row tile
    int ion[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int il = (w.tile * R + r) * 16 + cl;
        ion[r] = K.off + (il < K.n ? il : K.n - 1);
    }
    v4 acc[R][CT];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int t = 0; t < CT; t++) acc[r][t] = v4{0, 0, 0, 0};
    const bool kim = g & 1;
    const int ns = (P + 1) >> 1;
    struct Set {
        t2 Y[R], Z[R];
        WT Wv[CT];
    };
    // the lane's point p = 2 s + g / 2 as (iy, iz), advanced by 2 per load;
    // points past the box repeat the last one against zero W (kb_mf_wm)
    int lp = g >> 1, liy = 0, liz = lp;
    const WT *Bs = B;
    auto load = [&](Set &S) {
        const bool pok = lp < P;
        const int iy = pok ? liy : 2 * K.Ky, iz = pok ? liz : 2 * K.Kz;
        const int yo = (a.by - K.Ky + iy) * n, zo = (a.bz - K.Kz + iz) * n;
#pragma unroll
        for (int r = 0; r < R; r++) {
            S.Y[r] = T[yo + ion[r]];
            S.Z[r] = T[zo + ion[r]];
        }
#pragma unroll
        for (int t = 0; t < CT; t++) S.Wv[t] = Bs[t * 16];
        Bs += lp + 2 < 2 * ns ? 4 * NCp : 0;
        lp += 2;
        liz += 2;
        while (liz >= NZ) { liz -= NZ; liy++; }
    };
    auto compute = [&](const Set &S) {
        real av[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const real yr = (real)S.Y[r].x, yi = (real)S.Y[r].y, zr = (real)S.Z[r].x, zi = (real)S.Z[r].y;
            av[r] = kim ? yr * zi + yi * zr : yr * zr - yi * zi;
        }
#pragma unroll
        for (int t = 0; t < CT; t++) {
            const real bv = (real)S.Wv[t];
#pragma unroll
            for (int r = 0; r < R; r++) acc[r][t] = M::mma(av[r], bv, acc[r][t]);
        }
    };
    Set S0, S1;
    load(S0);
    int s = 0;
    for (; s + 1 < ns; s += 2) {
        load(S1);
        compute(S0);
        load(S0);
        compute(S1);
    }
    if (s < ns) compute(S0);
    // epilogue: lanes of even columns hold T_re, their neighbours T_im;
    // z = ex[nx] T per (ion, nx, m), summed with the nx factors into
    // M_b = sum_k n_b Im(W e^{ik.r}) and N_bd = sum_k n_b n_d Re(W e^{ik.r})
    const int nout = CH == 6 ? 9 : 3;
#pragma unroll
    for (int r = 0; r < R; r++) {
        double o[4][9];
#pragma unroll
        for (int e = 0; e < 4; e++)
#pragma unroll
            for (int c = 0; c < 9; c++) o[e][c] = 0.0;
#pragma unroll
        for (int t = 0; t < CT; t++) {
            const int col = t * 16 + cl;
            const bool use = col < NC && !(cl & 1);
            const int nx = col / (2 * CH), m = (col >> 1) % CH;
            const double fx = (double)nx;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const double tr = (double)acc[r][t][e];
                const double ti = __shfl_xor(tr, 1, 64);
                const int rl = (w.tile * R + r) * 16 + M::row(lane, e);
                const int io = K.off + (rl < K.n ? rl : K.n - 1);
                if (use) {
                    const t2 X = T[(nx < K.Kx ? nx : K.Kx) * n + io];
                    const double xr = (double)X.x, xi = (double)X.y;
                    const double zr = xr * tr - xi * ti, zi = xr * ti + xi * tr;
                    o[e][0] += m == 0 ? fx * zi : 0.0;
                    o[e][1] += m == 1 ? zi : 0.0;
                    o[e][2] += m == 2 ? zi : 0.0;
                    o[e][3] += m == 0 ? fx * fx * zr : 0.0;
                    o[e][4] += m == 1 ? fx * zr : 0.0;
                    o[e][5] += m == 2 ? fx * zr : 0.0;
                    o[e][6] += m == 3 ? zr : 0.0;
                    o[e][7] += m == 4 ? zr : 0.0;
                    o[e][8] += m == 5 ? zr : 0.0;
                }
            }
        }
#pragma unroll
        for (int e = 0; e < 4; e++)
#pragma unroll
            for (int c = 0; c < 9; c++) {
                if (c >= nout) break;
                double v = o[e][c];
                v += __shfl_xor(v, 1, 64);
                v += __shfl_xor(v, 2, 64);
                v += __shfl_xor(v, 4, 64);
                v += __shfl_xor(v, 8, 64);
                o[e][c] = v;
            }
        // lane cl = e < 4 of each lane group writes accumulator row e
#pragma unroll
        for (int e = 0; e < 4; e++) {
            if (cl != e) continue;
            const int rl = (w.tile * R + r) * 16 + M::row(lane, e);
            if (rl >= K.n) continue;
            const int io = K.off + rl;
            const double(*Bm)[3] = K.c.B;
            const double s1 = MDFF_TPI * K.tfpiV, s2 = MDFF_TPI * MDFF_TPI * K.tfpiV;
            double v[9];
            for (int x = 0; x < 3; x++) v[x] = s1 * (Bm[x][0] * o[e][0] + Bm[x][1] * o[e][1] + Bm[x][2] * o[e][2]);
            if (CH == 6) {
                const double Nm[3][3] = {{o[e][3], o[e][4], o[e][5]}, {o[e][4], o[e][6], o[e][7]}, {o[e][5], o[e][7], o[e][8]}};
                double BN[3][3];
                for (int x = 0; x < 3; x++)
                    for (int d = 0; d < 3; d++) BN[x][d] = Bm[x][0] * Nm[0][d] + Bm[x][1] * Nm[1][d] + Bm[x][2] * Nm[2][d];
                const int px[6] = {0, 0, 0, 1, 1, 2}, py[6] = {0, 1, 2, 1, 2, 2};
                for (int c = 0; c < 6; c++)
                    v[3 + c] = s2 * (BN[px[c]][0] * Bm[py[c]][0] + BN[px[c]][1] * Bm[py[c]][1] + BN[px[c]][2] * Bm[py[c]][2]);
            }
            if (a.ef) { a.ef[io] = v[0]; a.ef[n + io] = v[1]; a.ef[2 * (size_t)n + io] = v[2]; }
            if (CH == 6 && a.efg)
                for (int c = 0; c < 6; c++) a.efg[(size_t)c * n + io] = v[3 + c];
            if (a.f) {
                const double qi = a.q[io];
                double fx = qi * v[0], fy = qi * v[1], fz = qi * v[2];
                if (CH == 6) {
                    const double m0 = a.mu[io], m1 = a.mu[n + io], m2 = a.mu[2 * (size_t)n + io];
                    fx += v[3] * m0 + v[4] * m1 + v[5] * m2;
                    fy += v[4] * m0 + v[6] * m1 + v[7] * m2;
                    fz += v[5] * m0 + v[7] * m1 + v[8] * m2;
                }
                a.f[io] = fx; a.f[n + io] = fy; a.f[2 * (size_t)n + io] = fz;
            }
        }
    }
}

// ------------------------------------------------------- SCF, per config
// momentpolaSCF pim.c:236-278 per configuration: iteration 0 takes the
// predictor (the last converged dipoles, extrapolate_dipole_aspc at order 0
// with corrector weight 0), later iterations mu = pol^T E
__global__ void kb_scf_mu(int n, const int *__restrict__ cfg, const int *__restrict__ act, int first,
                          const int *__restrict__ type, const TypeTab *__restrict__ T, const double *__restrict__ hist,
                          const double *__restrict__ ef, double *__restrict__ mu, double *__restrict__ upol)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !act[cfg[i]]) return;
    int t = type[i];
    double m[3] = {0.0, 0.0, 0.0};
    if (first) {
        m[0] = hist[i]; m[1] = hist[n + i]; m[2] = hist[2 * n + i];
    } else if (T->polar[t]) {
        const double *P = T->pol[t];
        double e0 = ef[i], e1 = ef[n + i], e2 = ef[2 * n + i];
        for (int c = 0; c < 3; c++) m[c] = P[0 * 3 + c] * e0 + P[1 * 3 + c] * e1 + P[2 * 3 + c] * e2;
    }
    mu[i] = m[0]; mu[n + i] = m[1]; mu[2 * n + i] = m[2];
    double up = 0.0;
    if (T->polar[t]) {
        const double *I = T->invpol[t];
        for (int c = 0; c < 3; c++)
            for (int d = 0; d < 3; d++) up += m[c] * I[3 * c + d] * m[d];
    }
    upol[i] = 0.5 * up;
}

// ------------------------------- conjugate-gradient SCF of the mixed batch
// Per configuration, the fixed point mu = pol^T E(mu), E(mu) = E_stat + T mu,
// is the symmetric system (pol^-T - T) mu = E_stat (pol symmetric, T the
// dipole tensor with its self term): preconditioned by pol^T, with the field
// E and the residual r = E - pol^-T mu carried along (each step applies T
// to the search direction only).  Per-ion terms go to part[col][n] and are
// summed per configuration by kb_segsum; the stop test is get_rmsd_mu's
// (pim.c:431-449) on the carried field.

// (pol^-T x)_c = sum_d invpol[d][c] x_d; (pol^T x)_c = sum_d pol[d][c] x_d
__device__ __forceinline__ void cg_polT(const double *P, const double x[3], double y[3])
{
    for (int c = 0; c < 3; c++) y[c] = P[0 * 3 + c] * x[0] + P[1 * 3 + c] * x[1] + P[2 * 3 + c] * x[2];
}

__device__ __forceinline__ double cg_rmsd_ion(const double *P, const double m[3], const double e[3])
{
    double r = 0.0;
    for (int c = 0; c < 3; c++) {
        const double paa = P[4 * c];
        if (paa != 0.0) {
            const double d = m[c] / paa - e[c];
            r += d * d;
        }
    }
    return r;
}

// start: r = E - pol^-T mu, z = pol^T r, p = z; part[0] = r.z, part[1] the
// rmsd term of (mu, E) (adj = 0: the SCF) or |r|^2 (adj = 1: the adjoint
// solve, whose "field" is its right-hand side plus T lam)
__global__ void kb_cg_start(int n, const int *__restrict__ cfg, const int *__restrict__ act,
                            const int *__restrict__ type, const TypeTab *__restrict__ T, const double *__restrict__ ef,
                            const double *__restrict__ mu, double *__restrict__ r, double *__restrict__ z,
                            double *__restrict__ p, double *__restrict__ part, int adj)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double rz = 0.0, rm = 0.0;
    if (act[cfg[i]]) {
        const int t = type[i];
        double rv[3] = {0, 0, 0}, zv[3] = {0, 0, 0};
        if (T->polar[t]) {
            const double m[3] = {mu[i], mu[n + i], mu[2 * n + i]}, e[3] = {ef[i], ef[n + i], ef[2 * n + i]};
            double im[3];
            cg_polT(T->invpol[t], m, im);
            for (int c = 0; c < 3; c++) rv[c] = e[c] - im[c];
            cg_polT(T->pol[t], rv, zv);
            rz = rv[0] * zv[0] + rv[1] * zv[1] + rv[2] * zv[2];
            rm = adj ? rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2] : cg_rmsd_ion(T->pol[t], m, e);
        }
        for (int c = 0; c < 3; c++) {
            r[(size_t)c * n + i] = rv[c];
            z[(size_t)c * n + i] = zv[c];
            p[(size_t)c * n + i] = zv[c];
        }
    }
    part[i] = rz;
    part[n + i] = rm;
}

// q = pol^-T p - w (w = T p); part[0] = p.q
__global__ void kb_cg_pq(int n, const int *__restrict__ cfg, const int *__restrict__ act,
                         const int *__restrict__ type, const TypeTab *__restrict__ T, const double *__restrict__ p,
                         const double *__restrict__ w, double *__restrict__ q, double *__restrict__ part)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double pq = 0.0;
    if (act[cfg[i]]) {
        const int t = type[i];
        const double pv[3] = {p[i], p[n + i], p[2 * n + i]};
        double ip[3] = {0, 0, 0};
        if (T->polar[t]) cg_polT(T->invpol[t], pv, ip);
        for (int c = 0; c < 3; c++) {
            const double qc = ip[c] - w[(size_t)c * n + i];
            q[(size_t)c * n + i] = qc;
            pq += pv[c] * qc;
        }
    }
    part[i] = pq;
}

// alpha = rz / pq per configuration; mu += alpha p, r -= alpha q, E += alpha
// w, z = pol^T r; part[0] = r.z, part[1] as kb_cg_start's
__global__ void kb_cg_update(int n, const int *__restrict__ cfg, const int *__restrict__ act,
                             const int *__restrict__ type, const TypeTab *__restrict__ T,
                             const double *__restrict__ rz, const double *__restrict__ pqs,
                             const double *__restrict__ p, const double *__restrict__ q,
                             const double *__restrict__ w, double *__restrict__ mu, double *__restrict__ r,
                             double *__restrict__ ef, double *__restrict__ z, double *__restrict__ part, int adj)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c0 = cfg[i];
    double rzn = 0.0, rm = 0.0;
    if (act[c0]) {
        const double pq = pqs[c0];
        const double al = pq > 0.0 ? rz[c0] / pq : 0.0;
        const int t = type[i];
        double m[3], rv[3], e[3];
        for (int c = 0; c < 3; c++) {
            const size_t k = (size_t)c * n + i;
            m[c] = mu[k] + al * p[k];
            rv[c] = r[k] - al * q[k];
            e[c] = ef[k] + al * w[k];
            mu[k] = m[c];
            r[k] = rv[c];
            ef[k] = e[c];
        }
        double zv[3] = {0, 0, 0};
        if (T->polar[t]) {
            cg_polT(T->pol[t], rv, zv);
            rzn = rv[0] * zv[0] + rv[1] * zv[1] + rv[2] * zv[2];
            rm = adj ? rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2] : cg_rmsd_ion(T->pol[t], m, e);
        }
        for (int c = 0; c < 3; c++) z[(size_t)c * n + i] = zv[c];
    }
    part[i] = rzn;
    part[n + i] = rm;
}

// the loop condition of pim.c:236 per configuration on its rmsd (sums
// s[c][0] = r.z, s[c][1] = rmsd^2 npol), or with s2 (the adjoint solve)
// |r| <= tol |rhs| (s[c][1] = |r|^2, s2[c] = |rhs|^2); and beta = r.z /
// r.z_old; one block
__global__ void __launch_bounds__(1024) kb_cg_control(int nconf, const BConfD *__restrict__ conf,
                                                      const double *__restrict__ s, int first, int *act, int *iscf,
                                                      double *rmsd, double *rz, double *beta, int maxit, int minit,
                                                      double tol, int *any, const double *__restrict__ s2)
{
    __shared__ int sany;
    if (threadIdx.x == 0) sany = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < nconf; c += blockDim.x) {
        if (!act[c]) continue;
        const int it = iscf[c] + 1;
        const int np = conf[c].npol;
        const double rm = s2 ? (s2[c] > 0.0 ? sqrt(s[2 * c + 1] / s2[c]) : 0.0)
                             : sqrt(s[2 * c + 1] / (double)(np > 0 ? np : 1));
        const double rzn = s[2 * c];
        iscf[c] = it;
        rmsd[c] = rm;
        beta[c] = (!first && rz[c] > 0.0) ? rzn / rz[c] : 0.0;
        rz[c] = rzn;
        const int a = ((it < maxit && rm > tol) || it < minit) ? 1 : 0;
        act[c] = a;
        if (a) atomicOr(&sany, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) *any = sany;
}

// e = a + b (3n), the adjoint solve's carried "field" s + T lam
__global__ void kb_add3(size_t m, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ e)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < m) e[i] = a[i] + b[i];
}

// p = z + beta p on the active configurations
__global__ void kb_cg_dir(int n, const int *__restrict__ cfg, const int *__restrict__ act,
                          const double *__restrict__ beta, const double *__restrict__ z, double *__restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    const int c0 = cfg[i % n];
    if (!act[c0]) return;
    p[i] = z[i] + beta[c0] * p[i];
}

// u_pol = mu.pol^-1 mu / 2 of the final dipoles (pim.c:253,299)
__global__ void kb_cg_upol(int n, const int *__restrict__ type, const TypeTab *__restrict__ T,
                           const double *__restrict__ mu, double *__restrict__ upol)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = type[i];
    double up = 0.0;
    if (T->polar[t]) {
        const double m[3] = {mu[i], mu[n + i], mu[2 * n + i]};
        const double *I = T->invpol[t];
        for (int c = 0; c < 3; c++)
            for (int d = 0; d < 3; d++) up += m[c] * I[3 * c + d] * m[d];
    }
    upol[i] = 0.5 * up;
}

// get_rmsd_mu pim.c:431-449 per ion
__global__ void kb_scf_rmsd(int n, const int *__restrict__ cfg, const int *__restrict__ act,
                            const int *__restrict__ type, const TypeTab *__restrict__ T, const double *__restrict__ mu,
                            const double *__restrict__ ef, double *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !act[cfg[i]]) return;
    int t = type[i];
    double r = 0.0;
    if (T->polar[t]) {
        for (int c = 0; c < 3; c++) {
            double paa = T->pol[t][4 * c];
            if (paa != 0.0) {
                double e = mu[(size_t)c * n + i] / paa - ef[(size_t)c * n + i];
                r += e * e;
            }
        }
    }
    out[i] = r;
}

// loop condition of momentpolaSCF pim.c:236 for every configuration; one
// block; any = some configuration still iterating
__global__ void __launch_bounds__(1024) kb_scf_control(int nconf, const BConfD *__restrict__ conf,
                                                       const double *__restrict__ rmsd2, int *act, int *iscf,
                                                       double *rmsd, int maxit, int minit, double tol, int *any)
{
    __shared__ int sany;
    if (threadIdx.x == 0) sany = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < nconf; c += blockDim.x) {
        if (!act[c]) continue;
        int it = iscf[c] + 1;
        int np = conf[c].npol;
        double r = sqrt(rmsd2[c] / (double)(np > 0 ? np : 1));
        iscf[c] = it;
        rmsd[c] = r;
        int a = ((it < maxit && r > tol) || it < minit) ? 1 : 0;
        act[c] = a;
        if (a) atomicOr(&sany, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) *any = sany;
}

// ef = 2 selfa2 mu + dir + rec (+ base) on active configurations (ewald.c:235-270)
__global__ void kb_field_sum(int n, const int *__restrict__ cfg, const int *__restrict__ act,
                             const BConfD *__restrict__ conf, const double *mu, const double *ef_dir,
                             const double *ef_rec, const double *base, double *ef)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    int ion = i % n;
    int c = cfg[ion];
    if (act && !act[c]) return;
    double v = 2.0 * conf[c].selfa2 * mu[i] + ef_dir[i] + ef_rec[i];
    if (base) v += base[i];
    ef[i] = v;
}

__global__ void kb_set_conf_int(int nconf, int *p, int v)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nconf) p[c] = v;
}

// --------------------------------------------------------------- adjoint
// g = 2 (F - F_ref), per-ion |F - F_ref|^2
__global__ void kb_adj_g(int n, const double *__restrict__ f, const double *__restrict__ fx,
                         const double *__restrict__ fy, const double *__restrict__ fz, double *__restrict__ g,
                         double *__restrict__ li)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double dx = f[i] - fx[i], dy = f[n + i] - fy[i], dz = f[2 * n + i] - fz[i];
    g[i] = 2.0 * dx; g[n + i] = 2.0 * dy; g[2 * n + i] = 2.0 * dz;
    li[i] = dx * dx + dy * dy + dz * dz;
}

// derivative of the Tang-Toennies damping f(r) = 1 - c e^{-br} sum_k (br)^k/k!
// (order m) and of f' with respect to b and c:
//   f_b = c r e^{-x} x^m/m!,  f'_b = c e^{-x} x^m (1 - x + m)/m!,
//   f_c = -e^{-x} sum_k x^k/k!, f'_c = b e^{-x} x^m/m!,  x = b r
__device__ __forceinline__ void tt_param_derivs(const double *E, double b, double c, double r, int m, double &fb,
                                                double &fdb, double &fc, double &fdc)
{
    double x = b * r, ex = exp(-x);
    double sum = E[m];
    for (int k = m - 1; k >= 0; k--) sum = sum * x + E[k];
    double p = 1.0;
    for (int k = 0; k < m; k++) p *= x;
    double em = E[m] * p * ex;
    fb = c * r * em;
    fdb = c * em * (1.0 - x + (double)m);
    fc = -ex * sum;
    fdc = b * em;
}

struct AdjDirArgs {
    int n, cap;
    const BConfD *conf;
    const int *cfg;
    const double *s;
    const int *type;
    const unsigned *nbr;
    const int *cnt;
    const TypeTab *T;
    const double *q, *mu, *g, *lam;
    int lcouldamp, nt;
    double coul;
    double *out;    // kb_adj_dir: [3][n] D_g E (real space); kb_adj_damp: [2 nt][n]
};

// D_g E_k, real space, at the polarisable ions: for every source j in the
// Coulomb list, the change of the field at k from j when the ions move by g,
//   charge:  -q_j (T2 - T2D_kj) (g_k - g_j)        (damped as the field is)
//   dipole:  (T3 . mu_j) (g_k - g_j)               (both ions carry a dipole)
// with T2 = a2 r r - b2 I, T3 = -c3 rrr + a2 (delta r) of multipole_ES_dir.
template <int LANES>
__global__ void __launch_bounds__(256) kb_adj_dir(AdjDirArgs a)
{
    const int lane = threadIdx.x % LANES;
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / LANES;
    if (i >= a.n) return;
    const TypeTab &T = *a.T;
    const int n = a.n;
    const int ti = a.type[i];
    double o0 = 0, o1 = 0, o2 = 0;
    if (T.polar[ti]) {
        const BConfD &K = a.conf[a.cfg[i]];
        const double six = a.s[i], siy = a.s[n + i], siz = a.s[2 * n + i];
        const double mix = a.mu[i], miy = a.mu[n + i], miz = a.mu[2 * n + i];
        const bool dip_i = (mix != 0.0) || (miy != 0.0) || (miz != 0.0);
        const double gix = a.g[i], giy = a.g[n + i], giz = a.g[2 * n + i];
        const double al = K.alpha, al2 = al * al, al3 = al2 * al, al5 = al3 * al2;
        const double rpi = 1.0 / 1.7724538509055160273;
        const int m = a.cnt[i];
        const unsigned *nb = a.nbr + (size_t)i * a.cap;
        for (int e = lane; e < m; e += LANES) {
            unsigned ent = nb[e];
            if (!(ent & NBR_LR)) continue;
            int j = ent & NBR_IDX;
            double x, y, z;
            min_image(K.c, six - a.s[j], siy - a.s[n + j], siz - a.s[2 * n + j], x, y, z);
            double d2 = x * x + y * y + z * z;
            if (d2 > K.rc2l) continue;
            const int tj = a.type[j];
            const double qj = a.q[j];
            const double mjx = a.mu[j], mjy = a.mu[n + j], mjz = a.mu[2 * n + j];
            const bool dip_j = (mjx != 0.0) || (mjy != 0.0) || (mjz != 0.0);
            const double d = sqrt(d2), dm1 = 1.0 / d, dm3 = dm1 / d2, dm5 = dm3 / d2, dm7 = dm5 / d2;
            const double expon = exp(-al2 * d2) * rpi;
            const double F0 = erfc(al * d);
            const double F1 = F0 + 2.0 * al * d * expon;
            const double F2 = F1 + 4.0 * al3 * d2 * d * expon / 3.0;
            const double F3 = F2 + 8.0 * al5 * d2 * d2 * d * expon / 15.0;
            const double a2 = 3.0 * F2 * dm5, b2 = F1 * dm3, c3 = 15.0 * F3 * dm7;
            const double dgx = gix - a.g[j], dgy = giy - a.g[n + j], dgz = giz - a.g[2 * n + j];
            const double rdg = x * dgx + y * dgy + z * dgz;
            if (qj != 0.0) {
                double A2 = a2, B2 = b2;
                if (a.lcouldamp && T.ldamp[ti][tj]) {
                    double f, fd;
                    tt_damp<double>(T.E_TT, T.db[ti][tj], T.dc[ti][tj], d, f, fd, T.dk[ti][tj]);
                    double F1d = 1.0 - f, F2d = F1d + (d / 3.0) * fd;
                    A2 -= 3.0 * F2d * dm5;
                    B2 -= F1d * dm3;
                }
                o0 += -qj * (A2 * x * rdg - B2 * dgx);
                o1 += -qj * (A2 * y * rdg - B2 * dgy);
                o2 += -qj * (A2 * z * rdg - B2 * dgz);
            }
            if (dip_i && dip_j) {
                const double rmj = x * mjx + y * mjy + z * mjz;
                const double mdg = mjx * dgx + mjy * dgy + mjz * dgz;
                o0 += -c3 * x * rdg * rmj + a2 * (mjx * rdg + x * mdg + dgx * rmj);
                o1 += -c3 * y * rdg * rmj + a2 * (mjy * rdg + y * mdg + dgy * rmj);
                o2 += -c3 * z * rdg * rmj + a2 * (mjz * rdg + z * mdg + dgz * rmj);
            }
        }
    }
    o0 = group_sum<LANES>(o0);
    o1 = group_sum<LANES>(o1);
    o2 = group_sum<LANES>(o2);
    if (lane == 0) { a.out[i] = o0; a.out[n + i] = o1; a.out[2 * n + i] = o2; }
}

// gradient terms of the charge-dipole damping parameters: for the field at
// polarisable ion k from charge q_j (params [tk][tj]),
//   coul q_j [ mu_k . dT2D/dtheta . (g_k - g_j) + (lam_k . r) f_theta / d^3 ]
// with dT2D/dtheta = 3 dF2d r r / d^5 - dF1d I / d^3, dF1d = -f_theta,
// dF2d = -f_theta + d f'_theta / 3; out[(2 tj + {0 b, 1 c})][k]
template <int LANES>
__global__ void __launch_bounds__(256) kb_adj_damp(AdjDirArgs a)
{
    const int lane = threadIdx.x % LANES;
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / LANES;
    if (i >= a.n) return;
    const TypeTab &T = *a.T;
    const int n = a.n;
    const int ti = a.type[i];
    double acc[2 * NT];
    for (int t = 0; t < 2 * a.nt; t++) acc[t] = 0.0;
    if (T.polar[ti] && a.lcouldamp) {
        const BConfD &K = a.conf[a.cfg[i]];
        const double six = a.s[i], siy = a.s[n + i], siz = a.s[2 * n + i];
        const double mix = a.mu[i], miy = a.mu[n + i], miz = a.mu[2 * n + i];
        const double lx = a.lam[i], ly = a.lam[n + i], lz = a.lam[2 * n + i];
        const double gix = a.g[i], giy = a.g[n + i], giz = a.g[2 * n + i];
        const int m = a.cnt[i];
        const unsigned *nb = a.nbr + (size_t)i * a.cap;
        for (int e = lane; e < m; e += LANES) {
            unsigned ent = nb[e];
            if (!(ent & NBR_LR)) continue;
            int j = ent & NBR_IDX;
            const int tj = a.type[j];
            const double qj = a.q[j];
            if (qj == 0.0 || !T.ldamp[ti][tj]) continue;
            double x, y, z;
            min_image(K.c, six - a.s[j], siy - a.s[n + j], siz - a.s[2 * n + j], x, y, z);
            double d2 = x * x + y * y + z * z;
            if (d2 > K.rc2l) continue;
            const double d = sqrt(d2), dm3 = 1.0 / (d2 * d), dm5 = dm3 / d2;
            double fb, fdb, fc, fdc;
            tt_param_derivs(T.E_TT, T.db[ti][tj], T.dc[ti][tj], d, T.dk[ti][tj], fb, fdb, fc, fdc);
            const double dgx = gix - a.g[j], dgy = giy - a.g[n + j], dgz = giz - a.g[2 * n + j];
            const double rdg = x * dgx + y * dgy + z * dgz;
            const double mr = mix * x + miy * y + miz * z;
            const double mdg = mix * dgx + miy * dgy + miz * dgz;
            const double lr = lx * x + ly * y + lz * z;
            // theta = b
            double dF1 = -fb, dF2 = -fb + d * fdb / 3.0;
            double tb = 3.0 * dF2 * dm5 * mr * rdg - dF1 * dm3 * mdg + lr * fb * dm3;
            dF1 = -fc; dF2 = -fc + d * fdc / 3.0;
            double tc = 3.0 * dF2 * dm5 * mr * rdg - dF1 * dm3 * mdg + lr * fc * dm3;
            acc[2 * tj] += a.coul * qj * tb;
            acc[2 * tj + 1] += a.coul * qj * tc;
        }
    }
    for (int t = 0; t < 2 * a.nt; t++) {
        double v = group_sum<LANES>(acc[t]);
        if (lane == 0) a.out[(size_t)t * n + i] = v;
    }
}

struct AdjShortArgs {
    int n, cap, nt;
    const BConfD *conf;
    const int *cfg;
    const double *s;
    const int *type;
    const unsigned *nbr;
    const int *cnt;
    const TypeTab *T;
    const double *g;
    double *out;    // [4 nt][n]: d/dA, d/dB, d/dC, d/dD of ion i's pairs with type tj
};

// short-range part of dPhi/dtheta, Phi = sum g.F at fixed dipoles: for the
// BHMFTD pair force w(d) r (bhmftd.c:117-271), half of
// dw/dtheta (g_i - g_j).r per visit of the full list, with
//   dw/dA = B e^{-Bd}/d, dw/dB = A e^{-Bd}(1 - Bd)/d,
//   dw/dC = (f6'/d^6 - 6 f6/d^7)/d, dw/dD = (f8'/d^8 - 8 f8/d^9)/d
template <int LANES>
__global__ void __launch_bounds__(256) kb_adj_short(AdjShortArgs a)
{
    const int lane = threadIdx.x % LANES;
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / LANES;
    if (i >= a.n) return;
    const TypeTab &T = *a.T;
    const int n = a.n;
    const int ti = a.type[i];
    const BConfD &K = a.conf[a.cfg[i]];
    const double six = a.s[i], siy = a.s[n + i], siz = a.s[2 * n + i];
    const double gix = a.g[i], giy = a.g[n + i], giz = a.g[2 * n + i];
    double acc[4 * NT];
    for (int t = 0; t < 4 * a.nt; t++) acc[t] = 0.0;
    if (T.lbh) {
        const int m = a.cnt[i];
        const unsigned *nb = a.nbr + (size_t)i * a.cap;
        for (int e = lane; e < m; e += LANES) {
            unsigned ent = nb[e];
            if (!(ent & NBR_SR)) continue;
            int j = ent & NBR_IDX;
            int tj = a.type[j];
            if (T.A[ti][tj] == 0.0) continue;
            double x, y, z;
            min_image(K.c, six - a.s[j], siy - a.s[n + j], siz - a.s[2 * n + j], x, y, z);
            double d2 = x * x + y * y + z * z;
            if (!(d2 < K.rc2s)) continue;
            double d = sqrt(d2), id = 1.0 / d;
            double B = T.B[ti][tj];
            double eb = exp(-B * d);
            double f6 = 1, f8 = 1, fd6 = 0, fd8 = 0;
            if (T.lbhd) {
                tt_damp<double>(T.E_TT, T.BD[ti][tj], 1.0, d, f6, fd6, 6);
                tt_damp<double>(T.E_TT, T.BD[ti][tj], 1.0, d, f8, fd8, 8);
            }
            double ir2 = 1.0 / d2, ir6 = ir2 * ir2 * ir2, ir8 = ir6 * ir2;
            double proj = 0.5 * (x * (gix - a.g[j]) + y * (giy - a.g[n + j]) + z * (giz - a.g[2 * n + j]));
            acc[4 * tj + 0] += proj * B * eb * id;
            acc[4 * tj + 1] += proj * T.A[ti][tj] * eb * (1.0 - B * d) * id;
            acc[4 * tj + 2] += proj * (fd6 * ir6 - 6.0 * f6 * ir6 * id) * id;
            acc[4 * tj + 3] += proj * (fd8 * ir8 - 8.0 * f8 * ir8 * id) * id;
        }
    }
    for (int t = 0; t < 4 * a.nt; t++) {
        double v = group_sum<LANES>(acc[t]);
        if (lane == 0) a.out[(size_t)t * n + i] = v;
    }
}

// right-hand side of the adjoint: s_k = D_g E_k = (real space) + G_rec,k g_k
// + E_rec,k[delta rho], on polarisable ions (zero elsewhere); also |s|^2
__global__ void kb_adj_rhs(int n, const int *__restrict__ type, const TypeTab *__restrict__ T,
                           const double *__restrict__ sdir, const double *__restrict__ efg_rec,
                           const double *__restrict__ g, const double *__restrict__ e_delta, double *__restrict__ s,
                           double *__restrict__ s2)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[3] = {0.0, 0.0, 0.0};
    if (T->polar[type[i]]) {
        const double G[6] = {efg_rec[i], efg_rec[n + i], efg_rec[2 * (size_t)n + i], efg_rec[3 * (size_t)n + i],
                             efg_rec[4 * (size_t)n + i], efg_rec[5 * (size_t)n + i]};
        double g0 = g[i], g1 = g[n + i], g2 = g[2 * n + i];
        v[0] = sdir[i] + G[0] * g0 + G[1] * g1 + G[2] * g2 + e_delta[i];
        v[1] = sdir[n + i] + G[1] * g0 + G[3] * g1 + G[4] * g2 + e_delta[n + i];
        v[2] = sdir[2 * n + i] + G[2] * g0 + G[4] * g1 + G[5] * g2 + e_delta[2 * n + i];
    }
    s[i] = v[0]; s[n + i] = v[1]; s[2 * n + i] = v[2];
    s2[i] = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
}

// lam = pol s as the start of the adjoint iteration
__global__ void kb_adj_start(int n, const int *__restrict__ type, const TypeTab *__restrict__ T,
                             const double *__restrict__ s, double *__restrict__ lam)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int t = type[i];
    for (int c = 0; c < 3; c++) {
        double v = 0.0;
        if (T->polar[t]) {
            const double *P = T->pol[t];
            v = P[c * 3 + 0] * s[i] + P[c * 3 + 1] * s[n + i] + P[c * 3 + 2] * s[2 * n + i];
        }
        lam[(size_t)c * n + i] = v;
    }
}

// one Jacobi sweep of (pol^-1 - T) lam = s: residual r = pol^-1 lam - s - T lam
// (|r|^2 per ion), then lam = pol (s + T lam); eT = T lam
__global__ void kb_adj_sweep(int n, const int *__restrict__ cfg, const int *__restrict__ act,
                             const int *__restrict__ type, const TypeTab *__restrict__ T, const double *__restrict__ s,
                             const double *__restrict__ eT, double *__restrict__ lam, double *__restrict__ r2)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !act[cfg[i]]) return;
    int t = type[i];
    double rr = 0.0;
    if (T->polar[t]) {
        const double *P = T->pol[t], *I = T->invpol[t];
        double l[3] = {lam[i], lam[n + i], lam[2 * n + i]};
        double b[3] = {s[i] + eT[i], s[n + i] + eT[n + i], s[2 * n + i] + eT[2 * n + i]};
        for (int c = 0; c < 3; c++) {
            double v = I[3 * c] * l[0] + I[3 * c + 1] * l[1] + I[3 * c + 2] * l[2] - b[c];
            rr += v * v;
        }
        for (int c = 0; c < 3; c++) lam[(size_t)c * n + i] = P[c * 3 + 0] * b[0] + P[c * 3 + 1] * b[1] + P[c * 3 + 2] * b[2];
    }
    r2[i] = rr;
}

// convergence of the adjoint solve per configuration: |r| <= tol |s|
__global__ void __launch_bounds__(1024) kb_adj_control(int nconf, const double *__restrict__ r2,
                                                       const double *__restrict__ s2, int *act, int *iters,
                                                       double *resid, int maxit, double tol, int *any)
{
    __shared__ int sany;
    if (threadIdx.x == 0) sany = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < nconf; c += blockDim.x) {
        if (!act[c]) continue;
        int it = iters[c] + 1;
        iters[c] = it;
        double rel = s2[c] > 0.0 ? sqrt(r2[c] / s2[c]) : 0.0;
        resid[c] = rel;
        int a = (it < maxit && rel > tol) ? 1 : 0;
        act[c] = a;
        if (a) atomicOr(&sany, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) *any = sany;
}

// polarisability gradient per ion, out[9][n]: with R = pol^-T mu - E the
// SCF residual (mu = pol^T E, pim.c:637), d(pol^-T)/d pol_ab mu =
// -pol^-T e_b (pol^-T mu)_a, so dL/d pol_ab = coul (pol^-T mu)_a (pol^-1 lam)_b
__global__ void kb_adj_pol(int n, const int *__restrict__ type, const TypeTab *__restrict__ T, double coul,
                           const double *__restrict__ lam, const double *__restrict__ mu, double *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int t = type[i];
    double u[3] = {0, 0, 0}, w[3] = {0, 0, 0};
    if (T->polar[t]) {
        const double *I = T->invpol[t];
        double l[3] = {lam[i], lam[n + i], lam[2 * n + i]};
        double m[3] = {mu[i], mu[n + i], mu[2 * n + i]};
        for (int c = 0; c < 3; c++) {
            u[c] = I[3 * c] * l[0] + I[3 * c + 1] * l[1] + I[3 * c + 2] * l[2];   // (pol^-1 lam)_c
            w[c] = I[c] * m[0] + I[3 + c] * m[1] + I[6 + c] * m[2];               // (pol^-T mu)_c
        }
    }
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) out[(size_t)(3 * a + b) * n + i] = coul * w[a] * u[b];
}
