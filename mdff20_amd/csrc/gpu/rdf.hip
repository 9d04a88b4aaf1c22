// rdf.hip -- radial distribution histogram on the device, the work of get_gr
// (radial_distribution.c:180-231) over frames held in HBM: TRAJFF frames
// uploaded from the host, the engine's current positions, or the engine's
// device trajectory.
//
// Every ordered pair (i, j != i) of the reference has the same distance as
// (j, i) bit for bit (the fractional difference only changes sign, and
// nearbyint and the cartesian products are odd), so the kernel visits each
// unordered pair once: 256 x 256 tiles of ions, tile pairs I <= J, one block
// per tile pair and frame.  A pair within cutgr adds one to the block's LDS
// histogram under (bin, type of i, type of j); the ordered counts of the
// reference are U[b][a][c] + U[b][c][a] of that unordered histogram U.
// Arithmetic follows the reference's order of operations with no
// contraction (the file is compiled with -ffp-contract=off): kardir
// (cell.c:125-143), the difference, nearbyint (round half to even), the
// cartesian minimum image (pbc.c:7-14), r^2 summed x, y, z, the correctly
// rounded sqrt and division and the truncation (int)(r / resg), so the bins
// agree with the reference's for every pair.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "engine_view.h"
#include "mdff_host.h"
#include "mdff_rdf.h"

#define HIPCHK(x)                                                                                         \
    do {                                                                                                  \
        hipError_t e_ = (x);                                                                              \
        if (e_ != hipSuccess) {                                                                           \
            fprintf(stderr, "mdff: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            return MDFF_ERR_DEVICE;                                                                       \
        }                                                                                                 \
    } while (0)

namespace {

constexpr int RT = 256;                  // ions per tile (one per thread)
constexpr int RDF_LDS_MAX = 12288;       // histogram entries kept in LDS (48 KB beside the 7 KB tile)
constexpr double RDF_PI = 3.1415926535897932385;

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

struct RdfCell {
    double A[3][3], B[3][3];
};

// kardir of every ion of every frame (cell.c:132-134); with rt also the
// positions dirkar (cell.c:152-154) gives back from them, which get_gr leaves
// in the global arrays
__global__ void k_rdf_frac(int n, RdfCell c, const double *__restrict__ r, size_t rstride,
                           double *__restrict__ s, double *__restrict__ rt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t f = blockIdx.y;
    const double *rf = r + f * rstride;
    const double x = rf[i], y = rf[n + i], z = rf[2 * n + i];
    const double s0 = x * c.B[0][0] + y * c.B[0][1] + z * c.B[0][2];
    const double s1 = x * c.B[1][0] + y * c.B[1][1] + z * c.B[1][2];
    const double s2 = x * c.B[2][0] + y * c.B[2][1] + z * c.B[2][2];
    double *sf = s + f * 3 * (size_t)n;
    sf[i] = s0;
    sf[n + i] = s1;
    sf[2 * n + i] = s2;
    if (rt) {
        rt[i] = s0 * c.A[0][0] + s1 * c.A[1][0] + s2 * c.A[2][0];
        rt[n + i] = s0 * c.A[0][1] + s1 * c.A[1][1] + s2 * c.A[2][1];
        rt[2 * n + i] = s0 * c.A[0][2] + s1 * c.A[1][2] + s2 * c.A[2][2];
    }
}

struct RdfArgs {
    int n, ntype, nbins, ntiles, lds;
    double cut2, resg;
    RdfCell c;
    const double *s;          // fractional coordinates, 3 n per frame
    const int *type;
    unsigned long long *hist; // [nbins][ntype][ntype] unordered counts, then one slot of pairs out of range
};

// tile pair (I, J), I <= J, of linear index k (J-major upper triangle)
__device__ __forceinline__ void tile_pair(long long k, int &I, int &J)
{
    long long j = (long long)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
    while ((j + 1) * (j + 2) / 2 <= k) j++;
    while (j * (j + 1) / 2 > k) j--;
    J = (int)j;
    I = (int)(k - j * (j + 1) / 2);
}

template <bool ORTHO>
__global__ void __launch_bounds__(RT) k_rdf_pairs(RdfArgs a)
{
    extern __shared__ unsigned int lh[];
    __shared__ double sj[3][RT];
    __shared__ int tj[RT];
    const int n = a.n, nt = a.ntype;
    const int nh = a.nbins * nt * nt;
    int I, J;
    tile_pair(blockIdx.x, I, J);
    const double *s = a.s + (size_t)blockIdx.y * 3 * n;
    const int t = threadIdx.x;
    if (a.lds)
        for (int e = t; e < nh; e += RT) lh[e] = 0u;
    const int i = I * RT + t, jl = J * RT + t;
    sj[0][t] = jl < n ? s[jl] : 0.0;
    sj[1][t] = jl < n ? s[n + jl] : 0.0;
    sj[2][t] = jl < n ? s[2 * n + jl] : 0.0;
    tj[t] = jl < n ? a.type[jl] : -1;
    double xi = 0.0, yi = 0.0, zi = 0.0;
    int ti = -1;
    if (i < n) {
        xi = s[i];
        yi = s[n + i];
        zi = s[2 * n + i];
        ti = a.type[i];
    }
    __syncthreads();
    unsigned long long out = 0;
    if (ti >= 0) {
        // diagonal tiles: partners above i only
        const int j0 = (I == J) ? t + 1 : 0;
        const int jn = min(RT, n - J * RT);
        for (int jj = j0; jj < jn; jj++) {
            double dx = xi - sj[0][jj], dy = yi - sj[1][jj], dz = zi - sj[2][jj];
            dx = dx - rint(dx);
            dy = dy - rint(dy);
            dz = dz - rint(dz);
            double rx, ry, rz;
            if (ORTHO) {
                // the off-diagonal products of pbc.c:11-13 are zeros: adding
                // them leaves the sum unchanged (up to the sign of zero)
                rx = dx * a.c.A[0][0];
                ry = dy * a.c.A[1][1];
                rz = dz * a.c.A[2][2];
            } else {
                rx = dx * a.c.A[0][0] + dy * a.c.A[1][0] + dz * a.c.A[2][0];
                ry = dx * a.c.A[0][1] + dy * a.c.A[1][1] + dz * a.c.A[2][1];
                rz = dx * a.c.A[0][2] + dy * a.c.A[1][2] + dz * a.c.A[2][2];
            }
            const double r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < a.cut2) {
                const int b = (int)(sqrt(r2) / a.resg);
                if (b >= 0 && b < a.nbins) {
                    const int e = (b * nt + ti) * nt + tj[jj];
                    if (a.lds) atomicAdd(&lh[e], 1u);
                    else atomicAdd(&a.hist[e], 1ull);
                } else {
                    out++;
                }
            }
        }
    }
    if (out) atomicAdd(&a.hist[nh], out);
    if (!a.lds) return;
    __syncthreads();
    for (int e = t; e < nh; e += RT) {
        const unsigned int v = lh[e];
        if (v) atomicAdd(&a.hist[e], (unsigned long long)v);
    }
}

}  // namespace

struct mdff_rdf {
    int n = 0, ntype = 0, nbins = 0, device = 0;
    double resg = 0.0, cutgr = 0.0;
    hipStream_t st = nullptr;
    int *type = nullptr;                 // device
    double *r = nullptr, *s = nullptr;   // upload buffer and fractional coordinates (one frame)
    double *rt = nullptr;                // positions after kardir + dirkar
    double *sf = nullptr;                // fractional coordinates of trajectory frames
    size_t sf_frames = 0;
    unsigned long long *hist = nullptr;  // nbins ntype^2 + 1
    long long frames = 0;
    std::vector<int> nionit;
};

static void rdf_free(mdff_rdf *h)
{
    if (!h) return;
    if (h->st) (void)hipStreamSynchronize(h->st);
    void *p[] = {h->type, h->r, h->s, h->rt, h->sf, h->hist};
    for (void *q : p)
        if (q) (void)hipFree(q);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
}

static size_t hist_len(const mdff_rdf *h) { return (size_t)h->nbins * h->ntype * h->ntype + 1; }

static RdfCell cell_of(const double A[3][3])
{
    mdff_cell c;
    mdff_lattice(A, 1.0, &c);
    RdfCell rc;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            rc.A[i][j] = c.A[i][j];
            rc.B[i][j] = c.B[i][j];
        }
    return rc;
}

static bool ortho(const RdfCell &c)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (i != j && c.A[i][j] != 0.0) return false;
    return true;
}

// the histogram of nf frames of fractional coordinates s (3 n each)
static int rdf_pairs(mdff_rdf *h, const RdfCell &c, const double *s, const int *type, int nf, hipStream_t st)
{
    if (nf <= 0) return MDFF_OK;
    const int T = (int)cdiv(h->n, RT);
    const long long npair = (long long)T * (T + 1) / 2;
    if (npair > 0x7fffffffLL || nf > 65535) return MDFF_ERR_ARG;
    RdfArgs a;
    a.n = h->n;
    a.ntype = h->ntype;
    a.nbins = h->nbins;
    a.ntiles = T;
    const int nh = h->nbins * h->ntype * h->ntype;
    a.lds = nh <= RDF_LDS_MAX;
    a.cut2 = h->cutgr * h->cutgr;
    a.resg = h->resg;
    a.c = c;
    a.s = s;
    a.type = type;
    a.hist = h->hist;
    const size_t lds = a.lds ? sizeof(unsigned int) * (size_t)nh : 0;
    dim3 grid((unsigned)npair, (unsigned)nf);
    if (ortho(c)) hipLaunchKernelGGL(k_rdf_pairs<true>, grid, dim3(RT), lds, st, a);
    else hipLaunchKernelGGL(k_rdf_pairs<false>, grid, dim3(RT), lds, st, a);
    HIPCHK(hipGetLastError());
    h->frames += nf;
    return MDFF_OK;
}

extern "C" {

// default_rdf radial_distribution.c:37-42 on the widths of lattice
// (cell.c:43-47)
int mdff_rdf_defaults(const double A[3][3], double *cutgr, int *nbins)
{
    if (!A || !cutgr || !nbins) return MDFF_ERR_ARG;
    mdff_cell c;
    mdff_lattice(A, 1.0, &c);
    double w = c.w[0];
    if (c.w[1] < w) w = c.w[1];
    if (c.w[2] < w) w = c.w[2];
    *cutgr = 0.5 * w;
    *nbins = (int)(*cutgr / 0.1) + 1;
    return MDFF_OK;
}

mdff_rdf *mdff_rdf_create(int n, int ntype, const int *type, int nbins, double resg, double cutgr, int *status)
{
    int dummy;
    int *rc = status ? status : &dummy;
    *rc = MDFF_ERR_ARG;
    if (n < 2 || ntype < 1 || ntype > MDFF_NTYPEMAX || !type || nbins < 1 || !(resg > 0.0) || !(cutgr > 0.0))
        return nullptr;
    // every distance below cutgr has a bin
    if ((double)nbins < floor(cutgr / resg) + 1.0) return nullptr;
    if ((size_t)nbins * ntype * ntype > ((size_t)1 << 28)) return nullptr;
    std::vector<int> cnt(ntype, 0);
    for (int i = 0; i < n; i++) {
        if (type[i] < 0 || type[i] >= ntype) return nullptr;
        cnt[type[i]]++;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        *rc = MDFF_ERR_DEVICE;
        return nullptr;
    }
    mdff_rdf *h = new mdff_rdf;
    h->n = n;
    h->ntype = ntype;
    h->nbins = nbins;
    h->resg = resg;
    h->cutgr = cutgr;
    h->nionit = cnt;
    *rc = MDFF_ERR_NOMEM;
    if (hipGetDevice(&h->device) != hipSuccess || hipStreamCreate(&h->st) != hipSuccess ||
        hipMalloc((void **)&h->type, sizeof(int) * n) != hipSuccess ||
        hipMalloc((void **)&h->r, sizeof(double) * 3 * n) != hipSuccess ||
        hipMalloc((void **)&h->s, sizeof(double) * 3 * n) != hipSuccess ||
        hipMalloc((void **)&h->rt, sizeof(double) * 3 * n) != hipSuccess ||
        hipMalloc((void **)&h->hist, sizeof(unsigned long long) * hist_len(h)) != hipSuccess) {
        rdf_free(h);
        return nullptr;
    }
    if (hipMemcpy(h->type, type, sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(h->hist, 0, sizeof(unsigned long long) * hist_len(h)) != hipSuccess) {
        *rc = MDFF_ERR_DEVICE;
        rdf_free(h);
        return nullptr;
    }
    *rc = MDFF_OK;
    return h;
}

void mdff_rdf_destroy(mdff_rdf *h) { rdf_free(h); }

int mdff_rdf_reset(mdff_rdf *h)
{
    if (!h) return MDFF_ERR_ARG;
    HIPCHK(hipMemsetAsync(h->hist, 0, sizeof(unsigned long long) * hist_len(h), h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    h->frames = 0;
    return MDFF_OK;
}

int mdff_rdf_add_frame(mdff_rdf *h, const double A[3][3], double *rx, double *ry, double *rz, int write_back)
{
    if (!h || !A || !rx || !ry || !rz) return MDFF_ERR_ARG;
    const size_t n = (size_t)h->n;
    HIPCHK(hipMemcpyAsync(h->r, rx, sizeof(double) * n, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->r + n, ry, sizeof(double) * n, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->r + 2 * n, rz, sizeof(double) * n, hipMemcpyHostToDevice, h->st));
    const RdfCell c = cell_of(A);
    hipLaunchKernelGGL(k_rdf_frac, dim3(cdiv(n, 256), 1), dim3(256), 0, h->st, h->n, c, h->r, 3 * n, h->s,
                       write_back ? h->rt : nullptr);
    int rc = rdf_pairs(h, c, h->s, h->type, 1, h->st);
    if (rc) return rc;
    if (write_back) {
        HIPCHK(hipMemcpyAsync(rx, h->rt, sizeof(double) * n, hipMemcpyDeviceToHost, h->st));
        HIPCHK(hipMemcpyAsync(ry, h->rt + n, sizeof(double) * n, hipMemcpyDeviceToHost, h->st));
        HIPCHK(hipMemcpyAsync(rz, h->rt + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost, h->st));
    }
    HIPCHK(hipStreamSynchronize(h->st));
    return MDFF_OK;
}

// frames of an engine (its current positions, or every frame of its device
// trajectory), on the engine's stream: no host copy of the positions
static int engine_frames(mdff_rdf *h, mdff_ctx *ctx, bool traj)
{
    if (!h || !ctx) return MDFF_ERR_ARG;
    EngineView v;
    int rc = mdff_engine_view(ctx, &v);
    if (rc) return rc;
    if (v.n != h->n || v.ntype != h->ntype || v.device != h->device) return MDFF_ERR_ARG;
    const int nf = traj ? v.nframes : 1;
    if (nf == 0) return MDFF_OK;
    const size_t n = (size_t)h->n;
    // the engine's type of every ion must be the handle's (its counts
    // normalise g(r))
    {
        std::vector<int> te(n), th(n);
        HIPCHK(hipMemcpyAsync(te.data(), v.type, sizeof(int) * n, hipMemcpyDeviceToHost, v.st));
        HIPCHK(hipMemcpyAsync(th.data(), h->type, sizeof(int) * n, hipMemcpyDeviceToHost, h->st));
        HIPCHK(hipStreamSynchronize(v.st));
        HIPCHK(hipStreamSynchronize(h->st));
        if (te != th) return MDFF_ERR_ARG;
    }
    double *s = h->s;
    if (traj) {
        if (h->sf_frames < (size_t)nf) {
            if (h->sf) (void)hipFree(h->sf);
            h->sf = nullptr;
            h->sf_frames = 0;
            if (hipMalloc((void **)&h->sf, sizeof(double) * 3 * n * nf) != hipSuccess) return MDFF_ERR_NOMEM;
            h->sf_frames = nf;
        }
        s = h->sf;
    }
    RdfCell c;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            c.A[i][j] = v.A[i][j];
            c.B[i][j] = v.B[i][j];
        }
    hipLaunchKernelGGL(k_rdf_frac, dim3(cdiv(n, 256), nf), dim3(256), 0, v.st, h->n, c, traj ? v.frames : v.r,
                       traj ? v.frame_doubles : 3 * n, s, nullptr);
    if ((rc = rdf_pairs(h, c, s, h->type, nf, v.st))) return rc;
    HIPCHK(hipStreamSynchronize(v.st));
    return MDFF_OK;
}

int mdff_rdf_sample(mdff_rdf *h, mdff_ctx *ctx) { return engine_frames(h, ctx, false); }

// iters histograms of the engine's positions between two HIP events on the
// engine's stream (kardir and pair kernels); the counts are reset after
int mdff_rdf_time(mdff_rdf *h, mdff_ctx *ctx, int iters, double *ms)
{
    if (!h || !ctx || iters < 1 || !ms) return MDFF_ERR_ARG;
    EngineView v;
    int rc = mdff_engine_view(ctx, &v);
    if (rc) return rc;
    if (v.n != h->n || v.ntype != h->ntype || v.device != h->device) return MDFF_ERR_ARG;
    const size_t n = (size_t)h->n;
    RdfCell c;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            c.A[i][j] = v.A[i][j];
            c.B[i][j] = v.B[i][j];
        }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return MDFF_ERR_DEVICE;
    hipLaunchKernelGGL(k_rdf_frac, dim3(cdiv(n, 256), 1), dim3(256), 0, v.st, h->n, c, v.r, 3 * n, h->s, nullptr);
    rc = rdf_pairs(h, c, h->s, h->type, 1, v.st);
    if (!rc && hipEventRecord(e0, v.st) != hipSuccess) rc = MDFF_ERR_DEVICE;
    for (int k = 0; k < iters && !rc; k++) {
        hipLaunchKernelGGL(k_rdf_frac, dim3(cdiv(n, 256), 1), dim3(256), 0, v.st, h->n, c, v.r, 3 * n, h->s,
                           nullptr);
        rc = rdf_pairs(h, c, h->s, h->type, 1, v.st);
    }
    float t = 0.0f;
    if (!rc && (hipEventRecord(e1, v.st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t, e0, e1) != hipSuccess))
        rc = MDFF_ERR_DEVICE;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *ms = (double)t / iters;
    return mdff_rdf_reset(h);
}

int mdff_rdf_sample_trajectory(mdff_rdf *h, mdff_ctx *ctx) { return engine_frames(h, ctx, true); }

int mdff_rdf_counts(mdff_rdf *h, long long *counts, long long info[2])
{
    if (!h || !counts) return MDFF_ERR_ARG;
    std::vector<unsigned long long> u(hist_len(h));
    HIPCHK(hipMemcpyAsync(u.data(), h->hist, sizeof(unsigned long long) * u.size(), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    const int nt = h->ntype;
    for (int b = 0; b < h->nbins; b++)
        for (int x = 0; x < nt; x++)
            for (int y = 0; y < nt; y++) {
                const size_t k = ((size_t)b * nt + x) * nt + y, kt = ((size_t)b * nt + y) * nt + x;
                counts[k] = (long long)(u[k] + u[kt]);
            }
    if (info) {
        info[0] = h->frames;
        info[1] = 2 * (long long)u[hist_len(h) - 1];
    }
    return MDFF_OK;
}

// write_rdf radial_distribution.c:120-141 on ordered counts: the shell
// volume at the bin centre over omega, the count over nconf vol n_a n_b
int mdff_rdf_gr(const mdff_rdf *h, const long long *counts, long long nconf, double omega, double *r, double *g)
{
    if (!h || !counts || nconf < 1 || !(omega > 0.0) || !g) return MDFF_ERR_ARG;
    const int nt = h->ntype;
    const double resg = h->resg;
    for (int igr = 0; igr < h->nbins; igr++) {
        const double rr = ((double)igr + 0.5) * resg;
        double vol = 4 * RDF_PI * (resg * rr * rr + (pow(resg, 3)) / 12);
        vol = vol / omega;
        if (r) r[igr] = rr;
        for (int a = 0; a < nt; a++)
            for (int b = 0; b < nt; b++) {
                const size_t k = ((size_t)igr * nt + a) * nt + b;
                g[k] = ((double)counts[k]) / (double)((double)nconf * vol * h->nionit[a] * h->nionit[b]);
            }
    }
    return MDFF_OK;
}

}  // extern "C"
