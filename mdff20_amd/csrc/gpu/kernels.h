// kernels.h -- HIP kernels of the engine (included once, by engine.hip).
#pragma once
#include "dev.h"

// ---------------------------------------------------------------- reductions
// out[k] = sum_i in[k*ld + i], i < n; one 1024-thread block per column k,
// fixed summation order (bitwise reproducible).
__device__ __forceinline__ void colsum_block(const double *__restrict__ col, int n, double *__restrict__ out)
{
    __shared__ double sh[16];
    // two independent accumulators, then wave and block sums in a fixed order
    double s0 = 0.0, s1 = 0.0;
    int i = threadIdx.x;
    for (; i + 1024 < n; i += 2048) {
        s0 += col[i];
        s1 += col[i + 1024];
    }
    if (i < n) s0 += col[i];
    double s = s0 + s1;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x < 64) {
        double t = threadIdx.x < 16 ? sh[threadIdx.x] : 0.0;
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) t += __shfl_xor(t, o);
        if (threadIdx.x == 0) *out = t;
    }
}

__global__ void __launch_bounds__(1024) k_colsum(const double *__restrict__ in, int n, int ld,
                                                 double *__restrict__ out, const int *active)
{
    if (active && !*active) return;
    colsum_block(in + (size_t)blockIdx.x * ld, n, out + blockIdx.x);
}

// two k_colsum jobs in one launch: blocks 0..nc1-1 the first, the rest the
// second (same order of summation as k_colsum)
__global__ void __launch_bounds__(1024) k_colsum2(const double *__restrict__ in1, int n1, int ld1, int nc1,
                                                  double *__restrict__ out1, const double *__restrict__ in2, int n2,
                                                  int ld2, double *__restrict__ out2)
{
    const int b = blockIdx.x;
    if (b < nc1) colsum_block(in1 + (size_t)b * ld1, n1, out1 + b);
    else colsum_block(in2 + (size_t)(b - nc1) * ld2, n2, out2 + (b - nc1));
}

// ------------------------------------------------------------ fractional coords
// kardir cell.c:125-143 on a copy
__global__ void k_frac(int n, CellD c, const double *__restrict__ rx, const double *__restrict__ ry,
                       const double *__restrict__ rz, double *__restrict__ s)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = rx[i], y = ry[i], z = rz[i];
    s[i] = x * c.B[0][0] + y * c.B[0][1] + z * c.B[0][2];
    s[n + i] = x * c.B[1][0] + y * c.B[1][1] + z * c.B[1][2];
    s[2 * n + i] = x * c.B[2][0] + y * c.B[2][1] + z * c.B[2][2];
}

// ------------------------------------------------------------------ Verlet
// check_verletlist verlet.c:124-170: the two largest displacements since the
// last build.  One ion per thread, 256-thread blocks: each block's two
// largest go to part[2 b], part[2 b + 1]; the block that finishes last
// (atomic count in *cnt, reset by it) merges them and decides the rebuild.
// The pair is the two largest of the set whatever the merge order.
__device__ __forceinline__ void top2_merge(double &a1, double &a2, double b1, double b2)
{
    if (a1 >= b1) { a2 = fmax(a2, b1); }
    else { a2 = fmax(b2, a1); a1 = b1; }
}

__device__ __forceinline__ void top2_block(double &m1, double &m2, double *m1s, double *m2s)
{
    m1s[threadIdx.x] = m1;
    m2s[threadIdx.x] = m2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            double a1 = m1s[threadIdx.x], a2 = m2s[threadIdx.x];
            top2_merge(a1, a2, m1s[threadIdx.x + o], m2s[threadIdx.x + o]);
            m1s[threadIdx.x] = a1;
            m2s[threadIdx.x] = a2;
        }
        __syncthreads();
    }
    m1 = m1s[0];
    m2 = m2s[0];
}

__global__ void __launch_bounds__(256) k_vcheck(int n, CellD c, const double *__restrict__ s,
                                                const double *__restrict__ s0, double skin, int *ctl, int reset,
                                                double *part, unsigned *cnt)
{
    __shared__ double m1s[256], m2s[256];
    __shared__ int last;
    // reset: clear the overflow flag and the largest count of the build that
    // follows
    if (reset && blockIdx.x == 0 && threadIdx.x == 0) {
        ctl[C_OVERFLOW] = 0;
        ctl[C_MAXCNT] = 0;
    }
    const int nb = gridDim.x;
    double m1 = 0.0, m2 = 0.0;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        double x, y, z;
        min_image(c, s[i] - s0[i], s[n + i] - s0[n + i], s[2 * n + i] - s0[2 * n + i], x, y, z);
        m1 = sqrt(x * x + y * y + z * z);
    }
    top2_block(m1, m2, m1s, m2s);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = m1;
        part[2 * blockIdx.x + 1] = m2;
        __threadfence();
        last = atomicAdd(cnt, 1u) == (unsigned)(nb - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double a1 = 0.0, a2 = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) top2_merge(a1, a2, part[2 * b], part[2 * b + 1]);
    __syncthreads();
    top2_block(a1, a2, m1s, m2s);
    if (threadIdx.x == 0) {
        *cnt = 0u;
        int rebuild = ctl[C_FIRST] || (a1 + a2 > skin);
        ctl[C_REBUILD] = rebuild;
        if (rebuild) ctl[C_NUPD] += 1;
        ctl[C_FIRST] = 0;
    }
}

// full neighbour list, all pairs scanned (one wavefront per ion, ordered by j):
// entry j | NBR_SR if d <= rc_sr+skin | NBR_LR if d <= rc_lr+skin
// (gen_pbc_verletlist verlet.c:54-122 keeps each pair once; both orders are
// stored here so each ion owns its sums).  all=1 lists every j != i (no list,
// lverletL false: the kernels then apply only the cutoffs).
__global__ void __launch_bounds__(256) k_build_allpairs(int n, CellD c, const double *__restrict__ s,
                                                        double cut_sr2, double cut_lr2, int all, int cap,
                                                        unsigned *__restrict__ nbr, int *__restrict__ cnt,
                                                        double *__restrict__ s0, int *ctl)
{
    if (!ctl[C_REBUILD]) return;
    int lane = threadIdx.x & 63;
    int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    double six = s[i], siy = s[n + i], siz = s[2 * n + i];
    int count = 0;
    for (int j0 = 0; j0 < n; j0 += 64) {
        int j = j0 + lane;
        unsigned e = 0;
        if (j < n && j != i) {
            if (all) {
                e = (unsigned)j | NBR_SR | NBR_LR;
            } else {
                double x, y, z;
                min_image(c, six - s[j], siy - s[n + j], siz - s[2 * n + j], x, y, z);
                double d2 = x * x + y * y + z * z;
                if (d2 <= cut_sr2) e |= NBR_SR;
                if (d2 <= cut_lr2) e |= NBR_LR;
                if (e) e |= (unsigned)j;
            }
        }
        unsigned long long m = __ballot(e != 0);
        int pos = count + __popcll(m & ((1ull << lane) - 1ull));
        if (e != 0 && pos < cap) nbr[(size_t)i * cap + pos] = e;
        count += __popcll(m);
    }
    if (lane == 0) {
        cnt[i] = count < cap ? count : cap;
        if (count > cap) atomicMax(&ctl[C_OVERFLOW], count);
        atomicMax(&ctl[C_MAXCNT], count);
        s0[i] = six;
        s0[n + i] = siy;
        s0[2 * n + i] = siz;
    }
}

// --------------------------------------------------------- cell binning
// The list of gen_pbc_verletlist (verlet.c:54-122, an all-pairs scan there)
// from a cell grid: ions binned by fractional position into
// nc[0] x nc[1] x nc[2] cells, each cell's ions put in ascending index
// order, then one wavefront per ion scans the (2 reach + 1)^3 cells around
// its own.  The row of an ion lists its neighbours cell by cell in a fixed
// stencil order, ascending index within a cell, so the list (and every sum
// over it) is the same from run to run.
struct CellGrid {
    int nc[3], reach[3];  // cells per axis, stencil half-width per axis
    int ncell, nst;       // cells, stencil cells (<= 125)
    int *ccount;          // [ncell] ions per cell during binning (zero between builds)
    int *cstart;          // [ncell + 1] first ion of each cell in cidx
    int *cidx;            // [n] ions by cell, in arrival order
    int *csort;           // [n] ions by cell, ascending index within a cell
    int *cell;            // [n] cell of each ion
    int *slot;            // [n] arrival slot of each ion in its cell
};

__device__ __forceinline__ int cell_coord(double sv, int nc)
{
    const double u = sv - floor(sv);   // [0, 1): positions are never wrapped
    const int k = (int)(u * (double)nc);
    return k < 0 ? 0 : (k >= nc ? nc - 1 : k);
}

__global__ void __launch_bounds__(256) k_cell_bin(int n, const double *__restrict__ s, CellGrid g, const int *ctl)
{
    if (!ctl[C_REBUILD]) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int cx = cell_coord(s[i], g.nc[0]), cy = cell_coord(s[n + i], g.nc[1]), cz = cell_coord(s[2 * n + i], g.nc[2]);
    const int c = (cz * g.nc[1] + cy) * g.nc[0] + cx;
    g.cell[i] = c;
    g.slot[i] = atomicAdd(&g.ccount[c], 1);
}

// exclusive prefix of the cell counts (one block), which are cleared for
// the next binning
__global__ void __launch_bounds__(1024) k_cell_scan(CellGrid g, const int *ctl)
{
    if (!ctl[C_REBUILD]) return;
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (g.ncell + 1023) / 1024;
    const int c0 = min(t * per, g.ncell), c1 = min(c0 + per, g.ncell);
    int sum = 0;
    for (int c = c0; c < c1; c++) sum += g.ccount[c];
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - sum;
    for (int c = c0; c < c1; c++) {
        const int k = g.ccount[c];
        g.cstart[c] = run;
        g.ccount[c] = 0;
        run += k;
    }
    if (t == 1023) g.cstart[g.ncell] = part[1023];
}

__global__ void __launch_bounds__(256) k_cell_fill(int n, CellGrid g, const int *ctl)
{
    if (!ctl[C_REBUILD]) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    g.cidx[g.cstart[g.cell[i]] + g.slot[i]] = i;
}

// each cell's ions in ascending index order (the arrival order of the
// binning is not reproducible): ion i goes to the place of its rank among
// the ions of its cell, counted over the unordered list
__global__ void __launch_bounds__(256) k_cell_rank(int n, CellGrid g, const int *ctl)
{
    if (!ctl[C_REBUILD]) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int c = g.cell[i], b = g.cstart[c], e = g.cstart[c + 1];
    int rank = 0;
    for (int k = b; k < e; k++) rank += g.cidx[k] < i;
    g.csort[b + rank] = i;
}

// the rows from the cell grid: the entries of k_build_allpairs (j, NBR_SR
// within rc_sr + skin, NBR_LR within rc_lr + skin).  One block per home
// cell: the ions of the stencil around it are staged in LDS in stencil
// order (chunks of CB_CHUNK), and each wave scans them for the home ions it
// owns (home ion k goes to wave k % 4), appending in order with ballots.
#define CB_CHUNK 1024
#define CB_HOME 256
__global__ void __launch_bounds__(256) k_build_cells(int n, CellD c, const double *__restrict__ s, CellGrid g,
                                                    double cut_sr2, double cut_lr2, int cap,
                                                    unsigned *__restrict__ nbr, int *__restrict__ cnt,
                                                    double *__restrict__ s0, int *ctl)
{
    if (!ctl[C_REBUILD]) return;
    __shared__ int pre[129], cst[128];
    __shared__ double qx[CB_CHUNK], qy[CB_CHUNK], qz[CB_CHUNK];
    __shared__ int qj[CB_CHUNK];
    __shared__ int hcount[CB_HOME];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int home = blockIdx.x;
    const int cx = home % g.nc[0], cy = (home / g.nc[0]) % g.nc[1], cz = home / (g.nc[0] * g.nc[1]);
    const int wx = 2 * g.reach[0] + 1, wy = 2 * g.reach[1] + 1;
    // stencil prefix (wave 0: cells lane and lane + 64)
    if (w == 0) {
        int cntk[2], stk[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int k = lane + 64 * h;
            cntk[h] = 0;
            stk[h] = 0;
            if (k < g.nst) {
                const int dx = k % wx - g.reach[0], dy = (k / wx) % wy - g.reach[1], dz = k / (wx * wy) - g.reach[2];
                const int x = (cx + dx + g.nc[0]) % g.nc[0], y = (cy + dy + g.nc[1]) % g.nc[1];
                const int z = (cz + dz + g.nc[2]) % g.nc[2];
                const int cc = (z * g.nc[1] + y) * g.nc[0] + x;
                stk[h] = g.cstart[cc];
                cntk[h] = g.cstart[cc + 1] - stk[h];
            }
        }
        int incl0 = cntk[0], incl1 = cntk[1];
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int v0 = __shfl_up(incl0, o), v1 = __shfl_up(incl1, o);
            if (lane >= o) {
                incl0 += v0;
                incl1 += v1;
            }
        }
        const int tot0 = __shfl(incl0, 63);
        pre[lane] = incl0 - cntk[0];
        pre[lane + 64] = tot0 + incl1 - cntk[1];
        if (lane == 63) pre[128] = tot0 + incl1;
        cst[lane] = stk[0];
        cst[lane + 64] = stk[1];
    }
    __syncthreads();
    const int total = pre[128];
    const int hb = g.cstart[home], nh = g.cstart[home + 1] - hb;
    for (int h0 = 0; h0 < nh; h0 += CB_HOME) {
        const int hn = min(CB_HOME, nh - h0);
        for (int k = t; k < hn; k += 256) hcount[k] = 0;
        for (int k0 = 0; k0 < total; k0 += CB_CHUNK) {
            const int kn = min(CB_CHUNK, total - k0);
            __syncthreads();   // the previous chunk is scanned, hcount is zeroed
            for (int q = t; q < kn; q += 256) {
                const int k = k0 + q;
                int lo = 0, hi = g.nst - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (pre[mid] <= k) lo = mid;
                    else hi = mid - 1;
                }
                const int j = g.csort[cst[lo] + (k - pre[lo])];
                qj[q] = j;
                qx[q] = s[j];
                qy[q] = s[n + j];
                qz[q] = s[2 * n + j];
            }
            __syncthreads();
            for (int hk = w; hk < hn; hk += 4) {
                const int i = g.csort[hb + h0 + hk];
                const double six = s[i], siy = s[n + i], siz = s[2 * n + i];
                int count = hcount[hk];
                for (int q0 = 0; q0 < kn; q0 += 64) {
                    const int q = q0 + lane;
                    unsigned e = 0;
                    if (q < kn) {
                        const int j = qj[q];
                        if (j != i) {
                            double x, y, z;
                            min_image(c, six - qx[q], siy - qy[q], siz - qz[q], x, y, z);
                            const double d2 = x * x + y * y + z * z;
                            if (d2 <= cut_sr2) e |= NBR_SR;
                            if (d2 <= cut_lr2) e |= NBR_LR;
                            if (e) e |= (unsigned)j;
                        }
                    }
                    const unsigned long long m = __ballot(e != 0);
                    const int pos = count + __popcll(m & ((1ull << lane) - 1ull));
                    if (e != 0 && pos < cap) nbr[(size_t)i * cap + pos] = e;
                    count += __popcll(m);
                }
                if (lane == 0) hcount[hk] = count;
            }
        }
        __syncthreads();
        for (int hk = t; hk < hn; hk += 256) {
            const int i = g.csort[hb + h0 + hk];
            const int count = hcount[hk];
            cnt[i] = count < cap ? count : cap;
            if (count > cap) atomicMax(&ctl[C_OVERFLOW], count);
            atomicMax(&ctl[C_MAXCNT], count);
            s0[i] = s[i];
            s0[n + i] = s[n + i];
            s0[2 * n + i] = s[2 * n + i];
        }
        __syncthreads();
    }
}

// each row in ascending neighbour index, as the all-pairs scan leaves it
// (the pair kernels gather from monotone addresses, and both builds give
// the same rows): one wavefront per ion places every entry at its rank,
// counted over the row staged in LDS; rows longer than RS_MAX keep the
// cell order
#define RS_MAX 1024
__global__ void __launch_bounds__(256) k_row_sort(int n, int cap, unsigned *__restrict__ nbr,
                                                 const int *__restrict__ cnt, const int *ctl)
{
    if (!ctl[C_REBUILD]) return;
    __shared__ unsigned row[4][RS_MAX];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + w;
    if (i >= n) return;
    const int m = cnt[i];
    if (m > RS_MAX) return;
    unsigned *r = nbr + (size_t)i * cap;
    for (int p = lane; p < m; p += 64) row[w][p] = r[p];
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    for (int p = lane; p < m; p += 64) {
        const unsigned e = row[w][p], key = e & 0x3fffffffu;
        int rank = 0;
        for (int q = 0; q < m; q++) rank += (row[w][q] & 0x3fffffffu) < key;
        r[rank] = e;
    }
}

// ------------------------------------------------------------- short range
struct ShortArgs {
    int n, cap;
    CellD c;
    const double *s;
    const int *type;
    const unsigned *nbr;
    const int *cnt;
    const TypeTab *T;
    double rc2;
    double *f;            // 3n: force (overwritten)
    double *acc;          // 14n: per-ion halves u_bh, w_bh[6], u_lj, w_lj[6]
    const BConfD *conf;   // BATCH: per-configuration cell and cutoffs
    const int *cfg;       // BATCH: configuration of each ion
};

// engforce_bhmftd_pbc bhmftd.c:117-271 and engforce_nmlj_pbc nmlj.c:254-385,
// LANES lanes per ion over its full list.  BATCH: the cell and cutoff are
// those of the ion's configuration.
template <typename real, int LANES, bool BATCH = false>
__global__ void __launch_bounds__(256) k_short(ShortArgs a)
{
    const int lane = threadIdx.x % LANES;
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / LANES;
    if (i >= a.n) return;
    const CellD &cell = BATCH ? a.conf[a.cfg[i]].c : a.c;
    const double rc2 = BATCH ? a.conf[a.cfg[i]].rc2s : a.rc2;
    const TypeTab &T = *a.T;
    const int n = a.n;
    const int ti = a.type[i];
    const double six = a.s[i], siy = a.s[n + i], siz = a.s[2 * n + i];
    double fx = 0, fy = 0, fz = 0, ub = 0, ul = 0;
    double wb[6] = {0, 0, 0, 0, 0, 0}, wl[6] = {0, 0, 0, 0, 0, 0};
    const int m = a.cnt[i];
    const unsigned *nb = a.nbr + (size_t)i * a.cap;
    for (int e = lane; e < m; e += LANES) {
        unsigned ent = nb[e];
        if (!(ent & NBR_SR)) continue;
        int j = ent & NBR_IDX;
        int tj = a.type[j];
        double x, y, z;
        min_image(cell, six - a.s[j], siy - a.s[n + j], siz - a.s[2 * n + j], x, y, z);
        double d2 = x * x + y * y + z * z;
        if (!(d2 < rc2)) continue;
        if (T.lbh && T.A[ti][tj] != 0.0) {
            real rd2 = (real)d2;
            real ir2 = (real)1.0 / rd2, d = sqrt(rd2);
            // FP32: 1/d as a product (FP64 keeps the reference's divisions)
            const real id = sizeof(real) == 4 ? d * ir2 : (real)0;
            real B = (real)T.B[ti][tj];
            real erh = (real)T.A[ti][tj] * exp(-B * d);
            real ir6 = ir2 * ir2 * ir2;
            real ir8 = ir6 * ir2;
            ir6 *= (real)T.C[ti][tj];
            ir8 *= (real)T.D[ti][tj];
            real f6 = 1, f8 = 1, fd6 = 0, fd8 = 0;
            if (T.lbhd) {
                tt_damp<real>(T.E_TT, (real)T.BD[ti][tj], (real)1.0, d, f6, fd6, 6);
                tt_damp<real>(T.E_TT, (real)T.BD[ti][tj], (real)1.0, d, f8, fd8, 8);
            }
            real ir6d = ir6 * f6, ir8d = ir8 * f8;
            real ir7, ir9, wr;
            if (sizeof(real) == 4) {
                ir7 = (real)6.0 * ir6d * id;
                ir9 = (real)8.0 * ir8d * id;
                wr = (B * erh - ir7 - ir9 + ir6 * fd6 + ir8 * fd8) * id;
            } else {
                ir7 = (real)6.0 * ir6d / d;
                ir9 = (real)8.0 * ir8d / d;
                wr = (B * erh - ir7 - ir9 + ir6 * fd6 + ir8 * fd8) / d;
            }
            double w = (double)wr;
            ub += 0.5 * (double)(erh - ir6d - ir8d);
            double px = w * x, py = w * y, pz = w * z;
            fx += px; fy += py; fz += pz;
            wb[0] += 0.5 * x * px; wb[1] += 0.25 * (x * py + y * px); wb[2] += 0.25 * (x * pz + z * px);
            wb[3] += 0.5 * y * py; wb[4] += 0.25 * (y * pz + z * py); wb[5] += 0.5 * z * pz;
        }
        if (T.llj) {
            double sr2 = T.sigsq[ti][tj] / d2;
            double srp = pow(sr2, T.ptwo[ti][tj]);
            double srq = pow(sr2, T.qtwo[ti][tj]);
            double u = 0.0;
            if (T.trunc == MDFF_TRUNC_NO) u = T.epsqp[ti][tj] * (T.plj[ti][tj] * srq - T.qlj[ti][tj] * srp);
            else if (T.trunc == MDFF_TRUNC_LINEAR)
                u = T.epsqp[ti][tj] * (T.plj[ti][tj] * srq - T.qlj[ti][tj] * srp) - T.uc[ti][tj];
            ul += 0.5 * u;
            double w = T.fc[ti][tj] * (srq - srp) * sr2;
            double px = w * x, py = w * y, pz = w * z;
            fx += px; fy += py; fz += pz;
            wl[0] += 0.5 * x * px; wl[1] += 0.25 * (x * py + y * px); wl[2] += 0.25 * (x * pz + z * px);
            wl[3] += 0.5 * y * py; wl[4] += 0.25 * (y * pz + z * py); wl[5] += 0.5 * z * pz;
        }
    }
    fx = group_sum<LANES>(fx); fy = group_sum<LANES>(fy); fz = group_sum<LANES>(fz);
    ub = group_sum<LANES>(ub); ul = group_sum<LANES>(ul);
#pragma unroll
    for (int k = 0; k < 6; k++) { wb[k] = group_sum<LANES>(wb[k]); wl[k] = group_sum<LANES>(wl[k]); }
    if (lane == 0) {
        a.f[i] = fx; a.f[n + i] = fy; a.f[2 * n + i] = fz;
        a.acc[i] = ub;
        a.acc[(size_t)7 * n + i] = ul;
#pragma unroll
        for (int k = 0; k < 6; k++) { a.acc[(size_t)(1 + k) * n + i] = wb[k]; a.acc[(size_t)(8 + k) * n + i] = wl[k]; }
    }
}

// ------------------------------------------------------- Ewald real space
struct DirArgs {
    int n, cap;
    CellD c;
    const double *s;
    const int *type;
    const unsigned *nbr;
    const int *cnt;
    const TypeTab *T;
    const double *q;      // n
    const double *mu;     // 3n SoA
    double alpha, rc2;
    int lcouldamp;
    const int *active;    // when non-null, skip unless *active (SCF loop gate)
    double *ef;           // 3n (written when DO_EF)
    double *efg;          // 9n (written when EFG_Q||EFG_MM)
    double *f;            // 3n force, coulombic units excluded (written when DO_F)
    double *uw;           // 7n: energy halves (always) then 6 virial halves (DO_F)
    int q_zero;           // all charges zero: ions without dipole get nothing
    const BConfD *conf;   // BATCH: per-configuration cell, alpha and cutoff
    const int *cfg;       // BATCH: configuration of each ion; active is then per configuration
};

// multipole_ES_dir ewald.c:353-871 from ion i's side of its full list.
// Contributions to ion i of pair (i,j), r = r_i - r_j:
//   q-q   (ewald.c:703-731): u qi qj F0/d; ef_i -qj T1 (+qj T1D_i if damped);
//                            efg_i -qj T2; f_ij qi qj T1
//   mu-mu (ewald.c:737-761), only when both ions carry a dipole: u -mui.T2.muj;
//                            ef_i T2 muj; efg_i muj.T3; f_ij -mui.T3.muj
//   q-mu  (ewald.c:767-812): u -qi T1.muj + qj T1.mui (+ damped);
//                            f_ij -qi T2 muj + qj T2 mui (+ damped)
// with F_i -= f_ij, stress -(r f_ij + f_ij r) (ewald.c:815-830).
// BATCH: cell, alpha, cutoff and the active gate are per configuration.
template <typename real, int LANES, bool DO_EF, bool EFG_Q, bool EFG_MM, bool DO_F, bool BATCH = false>
__global__ void __launch_bounds__(256) k_ewald_dir(DirArgs a)
{
    if (!BATCH && a.active && !*a.active) return;
    const int lane = threadIdx.x % LANES;
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / LANES;
    if (i >= a.n) return;
    if (BATCH && a.active && !a.active[a.cfg[i]]) return;
    const CellD &cell = BATCH ? a.conf[a.cfg[i]].c : a.c;
    const double alpha = BATCH ? a.conf[a.cfg[i]].alpha : a.alpha;
    const double rc2 = BATCH ? a.conf[a.cfg[i]].rc2l : a.rc2;
    const TypeTab &T = *a.T;
    const int n = a.n;
    const int ti = a.type[i];
    const double six = a.s[i], siy = a.s[n + i], siz = a.s[2 * n + i];
    const real qi = (real)a.q[i];
    const real mix = (real)a.mu[i], miy = (real)a.mu[n + i], miz = (real)a.mu[2 * n + i];
    const bool dip_i = (a.mu[i] != 0.0) || (a.mu[n + i] != 0.0) || (a.mu[2 * n + i] != 0.0);
    const real al = (real)alpha, al2 = al * al, al3 = al2 * al, al5 = al3 * al2;
    const real rpi = (real)(1.0 / 1.7724538509055160273);
    double ex = 0, ey = 0, ez = 0;
    double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double fx = 0, fy = 0, fz = 0, uu = 0;
    double w6[6] = {0, 0, 0, 0, 0, 0};
    const int m = (a.q_zero && !dip_i) ? 0 : a.cnt[i];
    const unsigned *nb = a.nbr + (size_t)i * a.cap;
    for (int e = lane; e < m; e += LANES) {
        unsigned ent = nb[e];
        if (!(ent & NBR_LR)) continue;
        int j = ent & NBR_IDX;
        double dx, dy, dz;
        min_image(cell, six - a.s[j], siy - a.s[n + j], siz - a.s[2 * n + j], dx, dy, dz);
        double dd2 = dx * dx + dy * dy + dz * dz;
        if (dd2 > rc2) continue;
        const int tj = a.type[j];
        const real qj = (real)a.q[j];
        const double mjxd = a.mu[j], mjyd = a.mu[n + j], mjzd = a.mu[2 * n + j];
        const bool dip_j = (mjxd != 0.0) || (mjyd != 0.0) || (mjzd != 0.0);
        const real mjx = (real)mjxd, mjy = (real)mjyd, mjz = (real)mjzd;
        const real x = (real)dx, y = (real)dy, z = (real)dz, d2 = (real)dd2;
        const real d = sqrt(d2);
        // FP32: one reciprocal and products for the inverse powers and the
        // constant divisions (FP64 keeps the reference's divisions)
        constexpr bool R32 = sizeof(real) == 4;
        const real rd2 = R32 ? (real)1.0 / d2 : (real)0;
        const real dm1 = R32 ? d * rd2 : (real)1.0 / d;
        const real dm3 = R32 ? dm1 * rd2 : dm1 / d2;
        const real dm5 = R32 ? dm3 * rd2 : dm3 / d2;
        const real dm7 = R32 ? dm5 * rd2 : dm5 / d2;
        const real third = (real)(1.0 / 3.0), fifteenth = (real)(1.0 / 15.0);
        const bool ldamp = a.lcouldamp && T.ldamp[ti][tj];
        real f1 = 1, fd1 = 0, f2 = 1, fd2 = 0;
        if (ldamp) {
            tt_damp<real>(T.E_TT, (real)T.db[ti][tj], (real)T.dc[ti][tj], d, f1, fd1, T.dk[ti][tj]);
            tt_damp<real>(T.E_TT, (real)T.db[tj][ti], (real)T.dc[tj][ti], d, f2, fd2, T.dk[tj][ti]);
        }
        // screening functions ewald.c:524-532
        const real expon = exp(-al2 * d2) * rpi;
        const real F0 = fast_erfc(al * d);
        const real F1 = F0 + (real)2.0 * al * d * expon;
        const real d3 = d2 * d;
        const real F2 = F1 + (R32 ? (real)4.0 * al3 * d3 * expon * third : (real)4.0 * al3 * d3 * expon / (real)3.0);
        const real F3 =
            F2 + (R32 ? (real)8.0 * al5 * d3 * d2 * expon * fifteenth : (real)8.0 * al5 * d3 * d2 * expon / (real)15.0);
        const real F1d1 = (real)1.0 - f1, F1d2 = (real)1.0 - f2;
        const real d3rd = R32 ? d * third : d / (real)3.0;
        const real F2d1 = F1d1 + d3rd * fd1, F2d2 = F1d2 + d3rd * fd2;
        // T1 = -r F1/d^3 ; T2 = a2 r r - b2 I ; T3 = -c3 rrr + a2 (delta r)
        const real t1 = dm3 * F1, t1d1 = dm3 * F1d1, t1d2 = dm3 * F1d2;
        const real a2 = (real)3.0 * F2 * dm5, b2 = F1 * dm3;
        const real a2d1 = (real)3.0 * F2d1 * dm5, b2d1 = F1d1 * dm3;
        const real a2d2 = (real)3.0 * F2d2 * dm5, b2d2 = F1d2 * dm3;
        const real c3 = (real)15.0 * F3 * dm7;
        const real rmi = x * mix + y * miy + z * miz;
        const real rmj = x * mjx + y * mjy + z * mjz;
        real fij0 = 0, fij1 = 0, fij2 = 0;
        // charge-charge
        const bool qq = (qi != (real)0) || (qj != (real)0);
        if (qq) {
            uu += 0.5 * (double)(qi * qj * F0 * dm1);
            if (DO_EF) {
                // ef_i += -qj T1 (+ qj T1D1)
                real sc = qj * t1 - (ldamp ? qj * t1d1 : (real)0);
                ex += (double)(sc * x); ey += (double)(sc * y); ez += (double)(sc * z);
            }
            if (DO_F) {
                real sc = -qi * qj * t1;
                fij0 = sc * x; fij1 = sc * y; fij2 = sc * z;
            }
            if (EFG_Q) {
                // efg_i += -qj T2
                real A2 = -qj * a2, B2 = qj * b2;
                g[0] += (double)(A2 * x * x + B2); g[1] += (double)(A2 * x * y); g[2] += (double)(A2 * x * z);
                g[3] += (double)(A2 * y * x); g[4] += (double)(A2 * y * y + B2); g[5] += (double)(A2 * y * z);
                g[6] += (double)(A2 * z * x); g[7] += (double)(A2 * z * y); g[8] += (double)(A2 * z * z + B2);
            }
        }
        // dipole-dipole (both ions must carry a dipole)
        if (dip_i && dip_j) {
            const real mimj = mix * mjx + miy * mjy + miz * mjz;
            uu += 0.5 * (double)(-(a2 * rmi * rmj - b2 * mimj));
            if (DO_EF) {
                ex += (double)(a2 * x * rmj - b2 * mjx);
                ey += (double)(a2 * y * rmj - b2 * mjy);
                ez += (double)(a2 * z * rmj - b2 * mjz);
            }
            if (EFG_MM) {
                // efg_i[a][b] += sum_c muj_c T3[a][b][c]
                const real r_[3] = {x, y, z}, mj_[3] = {mjx, mjy, mjz};
#pragma unroll
                for (int p = 0; p < 3; p++)
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        real t = -c3 * r_[p] * r_[q] * rmj + a2 * (r_[q] * mj_[p] + r_[p] * mj_[q]);
                        if (p == q) t += a2 * rmj;
                        g[3 * p + q] += (double)t;
                    }
            }
            if (DO_F) {
                // f_ij,k = -sum mui_a T3[a][k][b] muj_b
                real cc = c3 * rmi * rmj - a2 * mimj;
                fij0 += cc * x - a2 * (mix * rmj + mjx * rmi);
                fij1 += cc * y - a2 * (miy * rmj + mjy * rmi);
                fij2 += cc * z - a2 * (miz * rmj + mjz * rmi);
            }
        }
        // charge-dipole
        if (dip_i || dip_j) {
            uu += 0.5 * (double)(qi * t1 * rmj - qj * t1 * rmi);
            if (ldamp) uu += 0.5 * (double)(-qi * t1d2 * rmj + qj * t1d1 * rmi);
            if (DO_F) {
                // -qi T2 muj + qj T2 mui  (+ -qj T2D1 mui + qi T2D2 muj)
                real cj = -qi * a2 * rmj, ci = qj * a2 * rmi;
                real lj = qi * b2, li = -qj * b2;
                if (ldamp) {
                    cj += qi * a2d2 * rmj; lj -= qi * b2d2;
                    ci -= qj * a2d1 * rmi; li += qj * b2d1;
                }
                fij0 += (cj + ci) * x + lj * mjx + li * mix;
                fij1 += (cj + ci) * y + lj * mjy + li * miy;
                fij2 += (cj + ci) * z + lj * mjz + li * miz;
            }
        }
        if (DO_F) {
            fx -= (double)fij0; fy -= (double)fij1; fz -= (double)fij2;
            double X = dx, Y = dy, Z = dz, f0 = fij0, f1_ = fij1, f2_ = fij2;
            w6[0] -= 0.5 * (X * f0 + X * f0);
            w6[1] -= 0.5 * (X * f1_ + Y * f0);
            w6[2] -= 0.5 * (X * f2_ + Z * f0);
            w6[3] -= 0.5 * (Y * f1_ + Y * f1_);
            w6[4] -= 0.5 * (Y * f2_ + Z * f1_);
            w6[5] -= 0.5 * (Z * f2_ + Z * f2_);
        }
    }
    if (DO_EF) { ex = group_sum<LANES>(ex); ey = group_sum<LANES>(ey); ez = group_sum<LANES>(ez); }
    if (EFG_Q || EFG_MM) {
#pragma unroll
        for (int k = 0; k < 9; k++) g[k] = group_sum<LANES>(g[k]);
    }
    uu = group_sum<LANES>(uu);
    if (DO_F) {
        fx = group_sum<LANES>(fx); fy = group_sum<LANES>(fy); fz = group_sum<LANES>(fz);
#pragma unroll
        for (int k = 0; k < 6; k++) w6[k] = group_sum<LANES>(w6[k]);
    }
    if (lane == 0) {
        if (DO_EF) { a.ef[i] = ex; a.ef[n + i] = ey; a.ef[2 * n + i] = ez; }
        if (EFG_Q || EFG_MM) {
#pragma unroll
            for (int k = 0; k < 9; k++) a.efg[(size_t)k * n + i] = g[k];
        }
        // the energy is formed in every call, as multipole_ES_dir does
        a.uw[i] = uu;
        if (DO_F) {
            a.f[i] = fx; a.f[n + i] = fy; a.f[2 * n + i] = fz;
#pragma unroll
            for (int k = 0; k < 6; k++) a.uw[(size_t)(1 + k) * n + i] = w6[k];
        }
    }
}

// ------------------------------------------------------------- k space
// Separable structure factor: e^{ik.r_j} = ex_j[nx] ey_j[ny] ez_j[nz] with
// k = 2 pi sum_b n_b B[.][b] (kspace.c:87-89) so k.r = 2 pi sum_b n_b t_b,
// t_b = sum_a B[a][b] r_a.  Tables hold n = 0..Kx for x and -K..K for y, z.
template <typename real>
__device__ __forceinline__ void eikr_store(real *t, size_t o, double cs, double sn)
{
    typedef real r2 __attribute__((ext_vector_type(2)));
    *(r2 *)(t + o) = r2{(real)cs, (real)sn};
}

template <typename real>
__global__ void k_eikr(int n, CellD c, int Kx, int Ky, int Kz, const double *__restrict__ rx,
                       const double *__restrict__ ry, const double *__restrict__ rz, const int *__restrict__ idx,
                       int nsel, real *__restrict__ tx, real *__restrict__ ty, real *__restrict__ tz,
                       const int *active, const int *__restrict__ pinv = nullptr, int npol = 0,
                       real *__restrict__ px = nullptr, real *__restrict__ py = nullptr, real *__restrict__ pz = nullptr)
{
    if (active && !*active) return;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsel) return;
    int i = idx ? idx[k] : k;
    // with px: the same rows also into the polarisable-ion tables (row pinv[i])
    const int pr = px ? pinv[i] : -1;
    double x = rx[i], y = ry[i], z = rz[i];
    double t[3];
    for (int b = 0; b < 3; b++) {
        double v = x * c.B[0][b] + y * c.B[1][b] + z * c.B[2][b];
        t[b] = v - rint(v);
    }
    // e^{i m th} for m = 0..K by the recurrence e^{i(m+1)th} = e^{i m th} e^{i th}
    // in FP64 (relative error below (m + 1) 2^-52), e^{-i m th} as its conjugate
    auto axis = [&](double th, int K, int off, real *tk, real *tp) {
        double s1, c1;
        sincos(th, &s1, &c1);
        double cs = 1.0, sn = 0.0;
        for (int m = 0; m <= K; m++) {
            eikr_store(tk, 2 * ((size_t)(off + m) * nsel + k), cs, sn);
            if (pr >= 0) eikr_store(tp, 2 * ((size_t)(off + m) * npol + pr), cs, sn);
            if (off > 0 && m > 0) {
                eikr_store(tk, 2 * ((size_t)(off - m) * nsel + k), cs, -sn);
                if (pr >= 0) eikr_store(tp, 2 * ((size_t)(off - m) * npol + pr), cs, -sn);
            }
            const double cn = cs * c1 - sn * s1, snn = sn * c1 + cs * s1;
            cs = cn;
            sn = snn;
        }
    };
    axis(MDFF_TPI * t[0], Kx, 0, tx, px);
    axis(MDFF_TPI * t[1], Ky, Ky, ty, py);
    axis(MDFF_TPI * t[2], Kz, Kz, tz, pz);
}

struct SkArgs {
    int nsel;                  // ions in the tables
    int Kx, Ky, Kz, P;         // P = (2Ky+1)(2Kz+1)
    const void *tx, *ty, *tz;  // tables (real)
    const int *idx;            // table row -> ion (nullptr = identity)
    const double *q, *mu;      // weights (q: n, mu: 3n SoA), full ion indexing
    int n;
    int chunk;                 // ions per block in y
    double *part;              // [nchunk][(Kx+1)*P][C][2]
    const int *active;
};

// forward S_c(k) = sum_j w_cj e^{ik.r_j}: block = (nx, 256 plane points),
// grid.y = ion chunks; partial sums per chunk (reduced in a fixed order).
// C = 1: q; 3: mu; 4: q and mu.
template <typename real, int C>
__global__ void __launch_bounds__(256) k_sk_fwd(SkArgs a)
{
    if (a.active && !*a.active) return;
    constexpr int SUB = 32;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int NY = 2 * a.Ky + 1, NZ = 2 * a.Kz + 1;
    real *shx = (real *)smem_raw;                 // SUB complex
    real *shy = shx + 2 * SUB;                    // NY*SUB complex
    real *shz = shy + 2 * NY * SUB;               // NZ*SUB complex
    real *shw = shz + 2 * NZ * SUB;               // C*SUB
    const int ntile = (a.P + 255) / 256;
    const int nx = blockIdx.x / ntile;
    const int p = (blockIdx.x % ntile) * 256 + threadIdx.x;
    const int j0 = blockIdx.y * a.chunk;
    const int j1 = min(j0 + a.chunk, a.nsel);
    const int iy = p / NZ, iz = p % NZ;
    const bool valid = p < a.P;
    const real *tx = (const real *)a.tx, *ty = (const real *)a.ty, *tz = (const real *)a.tz;
    double acc[C][2];
#pragma unroll
    for (int c = 0; c < C; c++) acc[c][0] = acc[c][1] = 0.0;
    for (int jb = j0; jb < j1; jb += SUB) {
        const int nj = min(SUB, j1 - jb);
        __syncthreads();
        for (int t = threadIdx.x; t < SUB; t += 256) {
            if (t < nj) {
                size_t o = 2 * ((size_t)nx * a.nsel + jb + t);
                shx[2 * t] = tx[o]; shx[2 * t + 1] = tx[o + 1];
                int ion = a.idx ? a.idx[jb + t] : jb + t;
                if (C == 1) shw[t] = (real)a.q[ion];
                if (C == 3) { shw[t] = (real)a.mu[ion]; shw[SUB + t] = (real)a.mu[a.n + ion]; shw[2 * SUB + t] = (real)a.mu[2 * a.n + ion]; }
                if (C == 4) {
                    shw[t] = (real)a.q[ion];
                    shw[SUB + t] = (real)a.mu[ion]; shw[2 * SUB + t] = (real)a.mu[a.n + ion]; shw[3 * SUB + t] = (real)a.mu[2 * a.n + ion];
                }
            }
        }
        for (int t = threadIdx.x; t < NY * SUB; t += 256) {
            int m = t / SUB, jj = t % SUB;
            if (jj < nj) {
                size_t o = 2 * ((size_t)m * a.nsel + jb + jj);
                shy[2 * t] = ty[o]; shy[2 * t + 1] = ty[o + 1];
            }
        }
        for (int t = threadIdx.x; t < NZ * SUB; t += 256) {
            int m = t / SUB, jj = t % SUB;
            if (jj < nj) {
                size_t o = 2 * ((size_t)m * a.nsel + jb + jj);
                shz[2 * t] = tz[o]; shz[2 * t + 1] = tz[o + 1];
            }
        }
        __syncthreads();
        if (valid) {
            real s[C][2];
#pragma unroll
            for (int c = 0; c < C; c++) s[c][0] = s[c][1] = (real)0;
            for (int jj = 0; jj < nj; jj++) {
                real xr = shx[2 * jj], xi = shx[2 * jj + 1];
                real yr = shy[2 * (iy * SUB + jj)], yi = shy[2 * (iy * SUB + jj) + 1];
                real zr = shz[2 * (iz * SUB + jj)], zi = shz[2 * (iz * SUB + jj) + 1];
                real pr = xr * yr - xi * yi, pi = xr * yi + xi * yr;
                real er = pr * zr - pi * zi, ei = pr * zi + pi * zr;
#pragma unroll
                for (int c = 0; c < C; c++) {
                    real w = shw[c * SUB + jj];
                    s[c][0] += w * er;
                    s[c][1] += w * ei;
                }
            }
#pragma unroll
            for (int c = 0; c < C; c++) { acc[c][0] += (double)s[c][0]; acc[c][1] += (double)s[c][1]; }
        }
    }
    if (valid) {
        size_t kf = (size_t)nx * a.P + p;
        size_t nkf = (size_t)(a.Kx + 1) * a.P;
        double *out = a.part + ((size_t)blockIdx.y * nkf + kf) * (2 * C);
#pragma unroll
        for (int c = 0; c < C; c++) { out[2 * c] = acc[c][0]; out[2 * c + 1] = acc[c][1]; }
    }
}

struct KRedArgs {
    int Kx, Ky, Kz, P, nchunk, C;
    const double *part;
    const double *kvec;   // [nkf][3]
    const double *Ak;     // [nkf], 0 outside the half space
    const double *kcoe;   // [nkf]
    double *W;            // [nkf][2] = Ak conj(rho)
    double *ek;           // [7][nkf]: str and 6 stress terms
    int do_stress;
    const int *active;
};

// rho(k) (kspace.c:189-302) from the channel sums, str = |rho|^2 Ak and the
// stress terms of multipole_ES_rec ewald.c:949-1006; W = Ak conj(rho).
// rho(k) from the channel sums S, then str, W and the stress terms
__device__ __forceinline__ void sk_reduce_S(const KRedArgs &a, size_t k, const double (&S)[4][2], double &Wr,
                                            double &Wi)
{
    size_t nkf = (size_t)(a.Kx + 1) * a.P;
    double Ak = a.Ak[k];
    if (Ak == 0.0) {
        a.W[2 * k] = 0.0; a.W[2 * k + 1] = 0.0;
        for (int t = 0; t < 7; t++) a.ek[t * nkf + k] = 0.0;
        Wr = Wi = 0.0;
        return;
    }
    double kx = a.kvec[3 * k], ky = a.kvec[3 * k + 1], kz = a.kvec[3 * k + 2];
    double rR = 0, rI = 0;
    int c0 = 0;
    if (a.C == 1 || a.C == 4) { rR += S[0][0]; rI += S[0][1]; c0 = 1; }
    if (a.C >= 3) {
        // + i (k.S_mu)
        double kr = kx * S[c0][0] + ky * S[c0 + 1][0] + kz * S[c0 + 2][0];
        double ki = kx * S[c0][1] + ky * S[c0 + 1][1] + kz * S[c0 + 2][1];
        rR += -ki;
        rI += kr;
    }
    double str = (rR * rR + rI * rI) * Ak;
    Wr = Ak * rR;
    Wi = -Ak * rI;
    a.W[2 * k] = Wr;
    a.W[2 * k + 1] = Wi;
    double *e = a.ek;
    e[k] = str;
    if (a.do_stress) {
        double kc = a.kcoe[k];
        e[1 * nkf + k] = (1.0 - kc * kx * kx) * str;
        e[2 * nkf + k] = -kc * kx * ky * str;
        e[3 * nkf + k] = -kc * kx * kz * str;
        e[4 * nkf + k] = (1.0 - kc * ky * ky) * str;
        e[5 * nkf + k] = -kc * ky * kz * str;
        e[6 * nkf + k] = (1.0 - kc * kz * kz) * str;
    } else {
        for (int t = 1; t < 7; t++) e[t * nkf + k] = 0.0;
    }
}

__device__ __forceinline__ void sk_reduce_one(const KRedArgs &a, size_t k, double &Wr, double &Wi)
{
    size_t nkf = (size_t)(a.Kx + 1) * a.P;
    double S[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    if (a.Ak[k] != 0.0) {
        for (int ch = 0; ch < a.nchunk; ch++) {
            const double *pp = a.part + ((size_t)ch * nkf + k) * (2 * a.C);
            for (int c = 0; c < a.C; c++) { S[c][0] += pp[2 * c]; S[c][1] += pp[2 * c + 1]; }
        }
    }
    sk_reduce_S(a, k, S, Wr, Wi);
}

__global__ void k_sk_reduce(KRedArgs a)
{
    if (a.active && !*a.active) return;
    size_t nkf = (size_t)(a.Kx + 1) * a.P;
    size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (k >= nkf) return;
    double Wr, Wi;
    sk_reduce_one(a, k, Wr, Wi);
}

struct BackArgs {
    int nsel, Kx, Ky, Kz, P;
    const void *tx, *ty, *tz;
    const double *W;        // [nkf][2]
    int rows_per_split;     // (nx,ny) rows per grid.y split
    double *part;           // [nsplit][NCH][nsel]
    const int *active;
};

// back-projection Z_j(k) = W_k e^{ik.r_j}; with k linear in n the sums over
// k of n_b Im Z and n_b n_c Re Z are formed per (nx,ny) row from the nz sums
//   V = sum_nz W ez, V1 = sum_nz nz W ez, V2 = sum_nz nz^2 W ez
// and the row factor P = ex[nx] ey[ny]:
//   H_x += nx Im(PV), H_y += ny Im(PV), H_z += Im(PV1)          (field)
//   M_xx += nx^2 Re(PV) ... M_zz += Re(PV2)                       (gradient)
// NCH = 3 (field) or 9 (field + gradient).
template <typename real, int NCH>
__global__ void __launch_bounds__(64) k_sk_back(BackArgs a)
{
    if (a.active && !*a.active) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int NZ = 2 * a.Kz + 1, NY = 2 * a.Ky + 1;
    real *shz = (real *)smem_raw;             // NZ*64 complex
    const int j = blockIdx.x * 64 + threadIdx.x;
    const bool valid = j < a.nsel;
    const real *tx = (const real *)a.tx, *ty = (const real *)a.ty, *tz = (const real *)a.tz;
    for (int m = 0; m < NZ; m++) {
        real zr = 0, zi = 0;
        if (valid) {
            size_t o = 2 * ((size_t)m * a.nsel + j);
            zr = tz[o]; zi = tz[o + 1];
        }
        shz[2 * (m * 64 + threadIdx.x)] = zr;
        shz[2 * (m * 64 + threadIdx.x) + 1] = zi;
    }
    __syncthreads();
    const int nrows = (a.Kx + 1) * NY;
    const int r0 = blockIdx.y * a.rows_per_split;
    const int r1 = min(r0 + a.rows_per_split, nrows);
    double H[3] = {0, 0, 0};
    double M[6] = {0, 0, 0, 0, 0, 0};
    for (int row = r0; row < r1; row++) {
        const int nx = row / NY, iy = row % NY, ny = iy - a.Ky;
        const double *Wr = a.W + 2 * ((size_t)nx * a.P + (size_t)iy * NZ);
        real v0r = 0, v0i = 0, v1r = 0, v1i = 0, v2r = 0, v2i = 0;
        for (int m = 0; m < NZ; m++) {
            const real wr = (real)Wr[2 * m], wi = (real)Wr[2 * m + 1];
            const real zr = shz[2 * (m * 64 + threadIdx.x)], zi = shz[2 * (m * 64 + threadIdx.x) + 1];
            const real ur = wr * zr - wi * zi, ui = wr * zi + wi * zr;
            const real nz = (real)(m - a.Kz);
            v0r += ur; v0i += ui;
            v1r += nz * ur; v1i += nz * ui;
            if (NCH == 9) { v2r += nz * nz * ur; v2i += nz * nz * ui; }
        }
        if (!valid) continue;
        size_t ox = 2 * ((size_t)nx * a.nsel + j), oy = 2 * ((size_t)iy * a.nsel + j);
        const real xr = tx[ox], xi = tx[ox + 1], yr = ty[oy], yi = ty[oy + 1];
        const double pr = (double)(xr * yr - xi * yi), pi = (double)(xr * yi + xi * yr);
        const double im0 = pr * v0i + pi * v0r, im1 = pr * v1i + pi * v1r;
        const double fnx = nx, fny = ny;
        H[0] += fnx * im0;
        H[1] += fny * im0;
        H[2] += im1;
        if (NCH == 9) {
            const double re0 = pr * v0r - pi * v0i, re1 = pr * v1r - pi * v1i, re2 = pr * v2r - pi * v2i;
            M[0] += fnx * fnx * re0;
            M[1] += fnx * fny * re0;
            M[2] += fnx * re1;
            M[3] += fny * fny * re0;
            M[4] += fny * re1;
            M[5] += re2;
        }
    }
    if (!valid) return;
    double *out = a.part + (size_t)blockIdx.y * NCH * a.nsel;
    out[j] = H[0];
    out[a.nsel + j] = H[1];
    out[2 * a.nsel + j] = H[2];
    if (NCH == 9)
        for (int t = 0; t < 6; t++) out[(size_t)(3 + t) * a.nsel + j] = M[t];
}

struct BackFinArgs {
    int n, nsel, nsplit, NCH;
    const int *idx;
    CellD c;
    double tfpiV;
    const double *part;
    const double *q, *mu;
    double *ef;       // 3n (DO_EF)
    double *efg;      // 9n (DO_EF: upper triangle + diagonal, ewald.c:966-973)
    double *f;        // 3n (DO_F)
    int do_ef, do_f;
    const int *active;
    // when ef_total is set: ef_total = 2 selfa2 mu_self + ef_dir + rec (+ base)
    // at the selected ions, the sum k_field_sum forms (ewald.c:235-270)
    double *ef_total;
    const double *mu_self, *ef_dir, *base;
    double selfa2;
    // with pq_part (and ef_total): per-block partial sums over the selected
    // ions of mu_self . (pol^-T mu_self - ef_total), the p.q of a
    // conjugate-gradient step (blocks of 256 selected ions)
    const int *type;
    const TypeTab *T;
    double *pq_part;
};

// sum the splits, map n-moments to k-moments and write field, gradient and
// force of multipole_ES_rec ewald.c:957-992 (times 8 pi / V, :1011-1034)
__device__ __forceinline__ double back_fin_one(const BackFinArgs &a, int t, const double *dir = nullptr)
{
    int i = a.idx ? a.idx[t] : t;
    double H[3] = {0, 0, 0}, M[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 4
    for (int s = 0; s < a.nsplit; s++) {
        const double *p = a.part + (size_t)s * a.NCH * a.nsel;
        for (int c = 0; c < 3; c++) H[c] += p[(size_t)c * a.nsel + t];
        if (a.NCH == 9)
            for (int c = 0; c < 6; c++) M[c] += p[(size_t)(3 + c) * a.nsel + t];
    }
    // E'_a = 2 pi sum_b B[a][b] H_b ; G_ab = (2 pi)^2 sum_cd B[a][c] B[b][d] M_cd
    double E[3];
    for (int aa = 0; aa < 3; aa++)
        E[aa] = MDFF_TPI * (a.c.B[aa][0] * H[0] + a.c.B[aa][1] * H[1] + a.c.B[aa][2] * H[2]);
    double Mf[3][3] = {{M[0], M[1], M[2]}, {M[1], M[3], M[4]}, {M[2], M[4], M[5]}};
    double G[3][3];
    for (int aa = 0; aa < 3; aa++)
        for (int bb = 0; bb < 3; bb++) {
            double sum = 0.0;
            for (int cc = 0; cc < 3; cc++)
                for (int dd = 0; dd < 3; dd++) sum += a.c.B[aa][cc] * a.c.B[bb][dd] * Mf[cc][dd];
            G[aa][bb] = MDFF_TPI * MDFF_TPI * sum;
        }
    const int n = a.n;
    double pq = 0.0;
    if (a.ef_total) {
        double v[3];
        for (int c = 0; c < 3; c++) {
            const size_t o = (size_t)c * n + i;
            v[c] = 2.0 * a.selfa2 * a.mu_self[o] + (dir ? dir[c] : a.ef_dir[o]) + a.tfpiV * E[c];
            if (a.base) v[c] += a.base[o];
            a.ef_total[o] = v[c];
        }
        if (a.pq_part) {
            const double *I = a.T->invpol[a.type[i]];
            const double p0 = a.mu_self[i], p1 = a.mu_self[n + i], p2 = a.mu_self[2 * n + i];
            const double pv[3] = {p0, p1, p2};
            for (int c = 0; c < 3; c++) pq += pv[c] * (I[c] * p0 + I[3 + c] * p1 + I[6 + c] * p2 - v[c]);
        }
    }
    if (a.do_ef) {
        a.ef[i] = a.tfpiV * E[0];
        a.ef[n + i] = a.tfpiV * E[1];
        a.ef[2 * n + i] = a.tfpiV * E[2];
        if (a.efg && a.NCH == 9) {
            a.efg[(size_t)0 * n + i] = a.tfpiV * G[0][0];
            a.efg[(size_t)1 * n + i] = a.tfpiV * G[0][1];
            a.efg[(size_t)2 * n + i] = a.tfpiV * G[0][2];
            a.efg[(size_t)3 * n + i] = 0.0;
            a.efg[(size_t)4 * n + i] = a.tfpiV * G[1][1];
            a.efg[(size_t)5 * n + i] = a.tfpiV * G[1][2];
            a.efg[(size_t)6 * n + i] = 0.0;
            a.efg[(size_t)7 * n + i] = 0.0;
            a.efg[(size_t)8 * n + i] = a.tfpiV * G[2][2];
        }
    }
    if (a.do_f) {
        double qi = a.q[i];
        double m0 = a.mu[i], m1 = a.mu[n + i], m2 = a.mu[2 * n + i];
        a.f[i] = a.tfpiV * (qi * E[0] + (a.NCH == 9 ? G[0][0] * m0 + G[0][1] * m1 + G[0][2] * m2 : 0.0));
        a.f[n + i] = a.tfpiV * (qi * E[1] + (a.NCH == 9 ? G[1][0] * m0 + G[1][1] * m1 + G[1][2] * m2 : 0.0));
        a.f[2 * n + i] = a.tfpiV * (qi * E[2] + (a.NCH == 9 ? G[2][0] * m0 + G[2][1] * m1 + G[2][2] * m2 : 0.0));
    }
    return pq;
}

__global__ void __launch_bounds__(256) k_sk_back_fin(BackFinArgs a)
{
    if (a.active && !*a.active) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const double pq = t < a.nsel ? back_fin_one(a, t) : 0.0;
    if (!a.pq_part) return;
    __shared__ double sh[256];
    sh[threadIdx.x] = pq;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.pq_part[blockIdx.x] = sh[0];
}

// One ion's T2 row: the field sum_j T2_ij mu_j over its m pair entries,
// lanes lane, lane + L, ... of L; U entries per lane per pass, so every
// index, tensor and dipole load of a pass is in flight together.
template <typename real, int L>
__device__ __forceinline__ void t2_row_sum(size_t row, int m, int lane, size_t plane, const int *__restrict__ idxp,
                                           const real *__restrict__ t2, const double *__restrict__ mu, int n,
                                           real &ex, real &ey, real &ez)
{
    constexpr int U = 4;
    for (int e0 = lane; e0 < m; e0 += U * L) {
        int j[U];
        real tt[U][6];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int e = e0 + u * L;
            j[u] = -1;
#pragma unroll
            for (int q = 0; q < 6; q++) tt[u][q] = 0;
            if (e < m) {
                const size_t o = row + e;
                j[u] = idxp[o];
#pragma unroll
                for (int q = 0; q < 6; q++) tt[u][q] = t2[q * plane + o];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (j[u] < 0) continue;
            const int jj = j[u];
            const real mx = (real)mu[jj], my = (real)mu[n + jj], mz = (real)mu[2 * n + jj];
            ex += tt[u][0] * mx + tt[u][1] * my + tt[u][2] * mz;
            ey += tt[u][1] * mx + tt[u][3] * my + tt[u][4] * mz;
            ez += tt[u][2] * mx + tt[u][4] * my + tt[u][5] * mz;
        }
    }
}

// k_sk_back_fin with the real-space field of the dipoles mu_self formed here
// from the pair tensors of k_t2_build (the sum of k_t2_field): L lanes per
// selected polarisable ion (selection = the polarisable rows in order) sum
// its tensor row, and lane 0 finishes the ion with that field in place of
// ef_dir.  pq partials per block of 256 / L ions.
template <typename real, int L>
__global__ void __launch_bounds__(256) k_sk_back_fin_t2(BackFinArgs a, int capP, int npol,
                                                        const int *__restrict__ cntp, const int *__restrict__ idxp,
                                                        const real *__restrict__ t2)
{
    if (a.active && !*a.active) return;
    const int lane = threadIdx.x % L;
    const int t = (blockIdx.x * blockDim.x + threadIdx.x) / L;
    const int n = a.n;
    const double *mu = a.mu_self;
    real ex = 0, ey = 0, ez = 0;
    if (t < a.nsel) {
        const int i = a.idx ? a.idx[t] : t;
        const bool dip_i = (mu[i] != 0.0) || (mu[n + i] != 0.0) || (mu[2 * n + i] != 0.0);
        if (dip_i) {
            const size_t plane = (size_t)npol * capP;
            const size_t row = (size_t)t * capP;
            const int m = cntp[t];
            t2_row_sum<real, L>(row, m, lane, plane, idxp, t2, mu, n, ex, ey, ez);
        }
    }
    const double d[3] = {group_sum<L>((double)ex), group_sum<L>((double)ey), group_sum<L>((double)ez)};
    const double pq = (t < a.nsel && lane == 0) ? back_fin_one(a, t, d) : 0.0;
    if (!a.pq_part) return;
    __shared__ double sh[256];
    sh[threadIdx.x] = pq;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.pq_part[blockIdx.x] = sh[0];
}

// ----------------------------------------------------------- elementwise
__global__ void k_zero(double *p, size_t n)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void k_zero_int(int *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// per-ion q^2 and |mu|^2 for the self energy (ewald.c:193-212)
__global__ void k_self_terms(int n, const double *q, const double *mu, double *out2n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double m0 = mu[i], m1 = mu[n + i], m2 = mu[2 * n + i];
    out2n[i] = q[i] * q[i];
    out2n[n + i] = m0 * m0 + m1 * m1 + m2 * m2;
}

// ef = self + dir + rec for the selected channels (ewald.c:235-270); the dir
// and rec arrays may be null when that part was not computed.
__global__ void k_field_sum(int n, double selfa2, const double *mu, const double *ef_dir, const double *ef_rec,
                            const double *base, double *ef, const int *active)
{
    if (active && !*active) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    double v = 2.0 * selfa2 * mu[i];
    if (ef_dir) v += ef_dir[i];
    if (ef_rec) v += ef_rec[i];
    if (base) v += base[i];
    ef[i] = v;
}

// ------------------------------------------ SCF real space over stored T2
// The dipole field of the SCF (multipole_ES_dir with charges off,
// ewald.c:737-761) needs, per pair of polarisable ions, only the screened
// tensor T2 = a2 r r - b2 I, which depends on the positions alone.  It is
// stored once per force call, per polarisable ion row (SoA: neighbour ion,
// then the 6 components xx xy xz yy yz zz), and each SCF evaluation is a
// gather sum E_i = sum_j T2_ij mu_j.
struct T2Args {
    int n, cap, capP;
    CellD c;
    const double *s;
    const unsigned *nbr;
    const int *cnt;
    const int *pidx;     // polarisable row -> ion
    const int *pinv;     // ion -> polarisable row or -1
    double alpha, rc2;
    int npol;
    int *cntp;           // [npol]
    int *idxp;           // [npol][capP] neighbour ion
    double *t2;          // [6][npol][capP]
    float *t2f;          // [6][npol][capP] (optional FP32 copy)
    int *overflow;
};

// one wavefront per polarisable ion, ballot-compacted rows
// real: the arithmetic of the tensor values (float when only t2f is wanted;
// the pair selection is FP64 either way)
template <typename real>
__global__ void __launch_bounds__(256) k_t2_build(T2Args a)
{
    const int lane = threadIdx.x & 63;
    const int pr = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pr >= a.npol) return;
    const int i = a.pidx[pr];
    const int n = a.n;
    const double six = a.s[i], siy = a.s[n + i], siz = a.s[2 * n + i];
    const double al = a.alpha, al2 = al * al, al3 = al2 * al;
    const double rpi = 1.0 / 1.7724538509055160273;
    const int m = a.cnt[i];
    const unsigned *nb = a.nbr + (size_t)i * a.cap;
    const size_t plane = (size_t)a.npol * a.capP;
    int count = 0;
    for (int e0 = 0; e0 < m; e0 += 64) {
        const int e = e0 + lane;
        bool keep = false;
        int j = 0;
        double T[6] = {0, 0, 0, 0, 0, 0};
        if (e < m) {
            unsigned ent = nb[e];
            j = ent & NBR_IDX;
            if ((ent & NBR_LR) && a.pinv[j] >= 0) {
                double x, y, z;
                min_image(a.c, six - a.s[j], siy - a.s[n + j], siz - a.s[2 * n + j], x, y, z);
                double d2 = x * x + y * y + z * z;
                if (d2 <= a.rc2) {
                    keep = true;
                    const real rx = (real)x, ry = (real)y, rz = (real)z, r2 = (real)d2;
                    const real ral = (real)al, ral2 = (real)al2, ral3 = (real)al3;
                    // FP32: one reciprocal and products (FP64: the reference's divisions)
                    constexpr bool R32 = sizeof(real) == 4;
                    const real d = sqrt(r2), ir2 = R32 ? (real)1 / r2 : (real)0;
                    const real dm1 = R32 ? d * ir2 : (real)1 / d;
                    const real dm3 = R32 ? dm1 * ir2 : dm1 / r2, dm5 = R32 ? dm3 * ir2 : dm3 / r2;
                    const real expon = exp(-ral2 * r2) * (real)rpi;
                    const real F0 = fast_erfc(ral * d);
                    const real F1 = F0 + (real)2 * ral * d * expon;
                    const real F2 = F1 + (R32 ? (real)4 * ral3 * r2 * d * expon * (real)(1.0 / 3.0)
                                              : (real)4 * ral3 * r2 * d * expon / (real)3);
                    const real a2 = (real)3 * F2 * dm5, b2 = F1 * dm3;
                    T[0] = a2 * rx * rx - b2; T[1] = a2 * rx * ry; T[2] = a2 * rx * rz;
                    T[3] = a2 * ry * ry - b2; T[4] = a2 * ry * rz; T[5] = a2 * rz * rz - b2;
                }
            }
        }
        unsigned long long bm = __ballot(keep);
        int pos = count + __popcll(bm & ((1ull << lane) - 1ull));
        if (keep && pos < a.capP) {
            const size_t o = (size_t)pr * a.capP + pos;
            a.idxp[o] = j;
            if (a.t2) {
#pragma unroll
                for (int k = 0; k < 6; k++) a.t2[k * plane + o] = T[k];
            }
            if (a.t2f) {
#pragma unroll
                for (int k = 0; k < 6; k++) a.t2f[k * plane + o] = (float)T[k];
            }
        }
        count += __popcll(bm);
    }
    if (lane == 0) {
        a.cntp[pr] = count < a.capP ? count : a.capP;
        if (count > a.capP) atomicMax(a.overflow, count);
    }
}

// E_i = sum_j T2_ij mu_j at every ion (zero where mu_i = 0 or i is not
// polarisable: the dipole-dipole field of the reference needs both ions to
// carry a dipole, ewald.c:737)
template <typename real, int LANES>
__global__ void __launch_bounds__(256) k_t2_field(int n, int capP, int npol, const int *__restrict__ pinv,
                                                   const int *__restrict__ cntp, const int *__restrict__ idxp,
                                                   const real *__restrict__ t2, const double *__restrict__ mu,
                                                   double *__restrict__ ef, const int *active)
{
    if (active && !*active) return;
    const int lane = threadIdx.x % LANES;
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / LANES;
    if (i >= n) return;
    const int pr = pinv[i];
    const bool dip_i = (mu[i] != 0.0) || (mu[n + i] != 0.0) || (mu[2 * n + i] != 0.0);
    real ex = 0, ey = 0, ez = 0;
    if (pr >= 0 && dip_i) {
        const size_t plane = (size_t)npol * capP;
        const size_t row = (size_t)pr * capP;
        const int m = cntp[pr];
        t2_row_sum<real, LANES>(row, m, lane, plane, idxp, t2, mu, n, ex, ey, ez);
    }
    double dx = group_sum<LANES>((double)ex), dy = group_sum<LANES>((double)ey), dz = group_sum<LANES>((double)ez);
    if (lane == 0) {
        ef[i] = dx;
        ef[n + i] = dy;
        ef[2 * n + i] = dz;
    }
}
