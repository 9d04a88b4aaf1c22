// batch.hip -- batched engine: many periodic configurations of uneven size
// evaluated together on one GPU (one launch per phase over the whole batch;
// the k-space phases are persistent grids on a work queue), and the adjoint
// of the force-matching loss.  Each configuration is evaluated as a static
// engforce (field.c:352-407) of the reference.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "batch_kernels.h"
#include "typetab.h"
// the engine's kernels, with internal linkage in this translation unit
namespace {
#include "kernels.h"
#include "md_kernels.h"
}  // namespace
#include "mdff_batch.h"
#include "mdff_gpu.h"
#include "mdff_host.h"

#define BCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            fprintf(stderr, "mdff batch: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            return MDFF_ERR_DEVICE;                                                               \
        }                                                                                         \
    } while (0)

namespace {

template <typename T>
int balloc(T **p, size_t n)
{
    if (n == 0) n = 1;
    if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) return MDFF_ERR_NOMEM;
    if (hipMemset(*p, 0, n * sizeof(T)) != hipSuccess) return MDFF_ERR_DEVICE;
    return MDFF_OK;
}

template <typename T>
void bfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

inline unsigned bdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

struct mdff_batch {
    mdff_params p;
    mdff_units u;
    int nconf = 0, n = 0, ntype = 0, precision = MDFF_FP64, device = 0, ncu = 256;
    hipStream_t st = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<BConfD> conf_h;
    std::vector<int> typ_h;
    TypeTab tab;
    int lpol = 0, ldmp = 0, lnonbonded = 0;
    // device arrays
    BConfD *conf = nullptr;
    TypeTab *tab_d = nullptr;
    int *cfg = nullptr, *type = nullptr;
    double *r = nullptr, *s = nullptr, *q = nullptr, *qz = nullptr, *mu = nullptr, *mu0 = nullptr, *hist = nullptr;
    double *ef_stat = nullptr, *ef = nullptr, *ef_dir = nullptr, *ef_rec = nullptr, *efg_rec = nullptr;
    double *efg_dummy = nullptr;
    double *f = nullptr, *f_sr = nullptr, *f_dir = nullptr, *f_rec = nullptr;
    double *acc_sr = nullptr, *uw_dir = nullptr, *selfb = nullptr, *upol_i = nullptr, *rmsd_i = nullptr;
    unsigned *nbr = nullptr;
    int *cnt = nullptr, *ctl = nullptr;
    int cap = 0;
    // k space
    KVec *kv = nullptr;
    size_t nktot = 0;
    double *W = nullptr, *ek = nullptr;
    double *tab64 = nullptr;
    float *tab32 = nullptr;
    // the same tables column by column ([TS][n]): the back-projection's
    // lanes (consecutive ions) read one column coalesced
    double *tabT64 = nullptr;
    float *tabT32 = nullptr;
    int TS = 0, by = 0, bz = 0;
    int *queue = nullptr;  // [2]: rho, [2]: back
    WItem *items_rho = nullptr, *items_back = nullptr;
    int nitems_rho = 0, nitems_back = 0;
    // matrix-core k-space (kb_mf_*): every configuration's full k box, the
    // box point -> k-vector map, the k-vector -> configuration map, the box
    // sums per k-vector and W on the boxes; work items of the two sums
    int mf = 0, mfr = 1;   // mfr: row tiles of a back-projection item (1 or 2)
    size_t nbox = 0, naw = 0, nwm = 0;
    std::vector<int> boxk_h;
    int *boxk = nullptr, *kcfg = nullptr;
    double *Sk = nullptr;
    double *aw = nullptr, *wm = nullptr;   // operand blocks (FP64, or FP32 in their first half)
    WItem *items_mf = nullptr, *items_mb = nullptr;
    int nitems_mf = 0, nitems_mb = 0;
    // items sorted by Kx (descending), then work: [begin, end) of each Kx
    std::vector<std::pair<int, int>> rng_mf, rng_mb;
    std::vector<int> kx_mf, kx_mb;
    // per-configuration control and sums
    int *act = nullptr, *iscf = nullptr, *any = nullptr, *any_h = nullptr;
    double *rmsd2 = nullptr, *rmsd = nullptr, *seg = nullptr, *seg_h = nullptr;
    // adjoint
    double *g = nullptr, *sdir = nullptr, *srhs = nullptr, *lam = nullptr, *eT = nullptr, *edel = nullptr;
    // conjugate-gradient SCF of the mixed form: r, z, p, w = T p, q = A p
    // [3][n]; per-ion terms [2][n]; per-configuration sums and scalars
    int cg = 0, cg_sym = 0;
    double *cg_e = nullptr;   // the adjoint solve's carried s + T lam
    double *cg_r = nullptr, *cg_z = nullptr, *cg_p = nullptr, *cg_w = nullptr, *cg_q = nullptr, *cg_part = nullptr;
    double *cg_s = nullptr, *cg_pq = nullptr, *cg_rz = nullptr, *cg_beta = nullptr;
    double *cols = nullptr, *tsum = nullptr;
    // tuning
    int lanes = 16, kpb = 64, ipb = 64, bpc = 4, tuned = 0;
    double last_ms = 0.0;
    long long sweeps = 0;
    std::vector<int> nscf_h;
    std::vector<double> rmsd_h;
};

static int lo_b(const mdff_batch *b) { return b->precision == MDFF_MIXED; }
static void cg_flags(mdff_batch *b);

// ------------------------------------------------------------ work items
static int build_items(mdff_batch *b)
{
    std::vector<WItem> ir, ib;
    std::vector<std::pair<long long, int>> order(b->nconf);
    for (int c = 0; c < b->nconf; c++) order[c] = {-(long long)b->conf_h[c].n * b->conf_h[c].nk, c};
    std::sort(order.begin(), order.end());  // largest configurations first
    for (auto &o : order) {
        const BConfD &K = b->conf_h[o.second];
        for (int t = 0; t * b->kpb < K.nk; t++) ir.push_back({o.second, t});
        for (int t = 0; t * b->ipb < K.n; t++) ib.push_back({o.second, t});
    }
    bfree(b->items_rho);
    bfree(b->items_back);
    int rc;
    if ((rc = balloc(&b->items_rho, ir.size())) || (rc = balloc(&b->items_back, ib.size()))) return rc;
    if (!ir.empty()) BCHK(hipMemcpy(b->items_rho, ir.data(), sizeof(WItem) * ir.size(), hipMemcpyHostToDevice));
    if (!ib.empty()) BCHK(hipMemcpy(b->items_back, ib.data(), sizeof(WItem) * ib.size(), hipMemcpyHostToDevice));
    b->nitems_rho = (int)ir.size();
    b->nitems_back = (int)ib.size();
    if (b->mf) {
        // per Kx (descending: the tile counts, and so the launch groups, fall
        // with Kx), largest work first: forward items ~ ions, back items ~
        // box points
        std::vector<std::pair<std::pair<int, long long>, WItem>> f, k;
        for (int c = 0; c < b->nconf; c++) {
            const BConfD &K = b->conf_h[c];
            const long long P = (long long)(2 * K.Ky + 1) * (2 * K.Kz + 1);
            for (int t = 0; t * 32 < P; t++) f.push_back({{-K.Kx, -(long long)K.n}, {c, t}});
            for (int t = 0; t * 16 * b->mfr < K.n; t++) k.push_back({{-K.Kx, -P}, {c, t}});
        }
        auto byw = [](const std::pair<std::pair<int, long long>, WItem> &x,
                      const std::pair<std::pair<int, long long>, WItem> &y) {
            if (x.first != y.first) return x.first < y.first;
            return x.second.conf != y.second.conf ? x.second.conf < y.second.conf : x.second.tile < y.second.tile;
        };
        std::sort(f.begin(), f.end(), byw);
        std::sort(k.begin(), k.end(), byw);
        auto ranges = [](const std::vector<std::pair<std::pair<int, long long>, WItem>> &v,
                         std::vector<std::pair<int, int>> &rng, std::vector<int> &kx, std::vector<WItem> &out) {
            rng.clear(); kx.clear(); out.clear();
            for (size_t i = 0; i < v.size(); i++) {
                if (i == 0 || v[i].first.first != v[i - 1].first.first) {
                    rng.push_back({(int)i, (int)i});
                    kx.push_back(-v[i].first.first);
                }
                rng.back().second = (int)i + 1;
                out.push_back(v[i].second);
            }
        };
        std::vector<WItem> fi, ki;
        ranges(f, b->rng_mf, b->kx_mf, fi);
        ranges(k, b->rng_mb, b->kx_mb, ki);
        bfree(b->items_mf);
        bfree(b->items_mb);
        if ((rc = balloc(&b->items_mf, fi.size())) || (rc = balloc(&b->items_mb, ki.size()))) return rc;
        if (!fi.empty()) BCHK(hipMemcpy(b->items_mf, fi.data(), sizeof(WItem) * fi.size(), hipMemcpyHostToDevice));
        if (!ki.empty()) BCHK(hipMemcpy(b->items_mb, ki.data(), sizeof(WItem) * ki.size(), hipMemcpyHostToDevice));
        b->nitems_mf = (int)fi.size();
        b->nitems_mb = (int)ki.size();
    }
    return MDFF_OK;
}

static void segsum(mdff_batch *b, const double *in, size_t ld, int ncol, int kspace, double *out)
{
    hipLaunchKernelGGL(kb_segsum, dim3(b->nconf, ncol), dim3(256), 0, b->st, in, ld, ncol, b->conf, kspace, out);
}

// ------------------------------------------------------------ launches
template <typename real, int MODE>
static void rho_t(mdff_batch *b, RhoArgs &a)
{
    int grid = std::max(1, std::min(a.nitems, b->bpc * b->ncu));
    switch (b->kpb) {
    case 32: hipLaunchKernelGGL((kb_rho<real, MODE, 32>), dim3(grid), dim3(256), 0, b->st, a); break;
    case 128: hipLaunchKernelGGL((kb_rho<real, MODE, 128>), dim3(grid), dim3(256), 0, b->st, a); break;
    default: hipLaunchKernelGGL((kb_rho<real, MODE, 64>), dim3(grid), dim3(256), 0, b->st, a); break;
    }
}

// rho(k), W and (when ek) the energy/stress terms of every active configuration
static void launch_rho(mdff_batch *b, int lo, int mode, const double *q, const double *mu, const double *g,
                       int do_ek, int do_stress, const int *act, bool exact = false)
{
    if (b->mf && mode != RM_ADJ) {
        // matrix-core sums over the k boxes, then rho, W, energy per k-vector.
        // The mixed form's SCF fields (exact) keep FP64 arithmetic on its FP32
        // tables: the field stays linear in the dipoles to FP64 rounding, as
        // the Jacobi stop test at conv_tol_ind needs; its static field and
        // force call use FP32 matrix instructions.
        const int C = mode == RM_Q ? 1 : (mode == RM_MU ? 3 : 4);
        const int form = !lo ? 0 : (exact ? 1 : 2);   // (tables, A, arithmetic): ddd, fdd, fff
        const void *tab = lo ? (const void *)b->tabT32 : (const void *)b->tabT64;
        const dim3 ga(16, b->nconf);
        if (form == 0) hipLaunchKernelGGL((kb_mf_aw<double, double>), ga, dim3(256), 0, b->st, b->conf, tab, b->n, C, q, mu, b->aw, act);
        else if (form == 1) hipLaunchKernelGGL((kb_mf_aw<float, double>), ga, dim3(256), 0, b->st, b->conf, tab, b->n, C, q, mu, b->aw, act);
        else hipLaunchKernelGGL((kb_mf_aw<float, float>), ga, dim3(256), 0, b->st, b->conf, tab, b->n, C, q, mu, (float *)b->aw, act);
        MFwdArgs f;
        f.conf = b->conf; f.tabT = tab; f.aw = b->aw;
        f.by = b->by; f.bz = b->bz; f.n = b->n; f.C = C;
        f.boxk = b->boxk; f.Sk = b->Sk; f.act = act;
        // one launch per run of Kx values with the same row-tile count
        for (size_t gi = 0; gi < b->rng_mf.size();) {
            const int MT = bmf_mt(b->kx_mf[gi], C);
            size_t gj = gi;
            while (gj + 1 < b->rng_mf.size() && bmf_mt(b->kx_mf[gj + 1], C) == MT) gj++;
            f.items = b->items_mf + b->rng_mf[gi].first;
            f.nitems = b->rng_mf[gj].second - b->rng_mf[gi].first;
            const dim3 gf(bdiv(f.nitems, 4));
#define MFWD(TT, AT, RT, M) hipLaunchKernelGGL((kb_mf_fwd<TT, AT, RT, M, 2>), gf, dim3(256), 0, b->st, f)
#define MFWD_T(TT, AT, RT) do { if (MT <= 2) MFWD(TT, AT, RT, 2); else if (MT <= 4) MFWD(TT, AT, RT, 4); \
                               else if (MT <= 6) MFWD(TT, AT, RT, 6); else MFWD(TT, AT, RT, 8); } while (0)
            if (form == 0) MFWD_T(double, double, double);
            else if (form == 1) MFWD_T(float, double, double);
            else MFWD_T(float, float, float);
#undef MFWD_T
#undef MFWD
            gi = gj + 1;
        }
        hipLaunchKernelGGL(kb_mf_rho, dim3(bdiv(b->nktot, 256)), dim3(256), 0, b->st, (int)b->nktot, b->kcfg, b->kv,
                           b->Sk, C, b->W, do_ek ? b->ek : nullptr, do_stress, act);
        return;
    }
    RhoArgs a;
    a.conf = b->conf; a.kv = b->kv;
    a.tab = lo ? (const void *)b->tab32 : (const void *)b->tab64;
    a.TS = b->TS; a.by = b->by; a.bz = b->bz; a.n = b->n;
    a.q = q; a.mu = mu; a.g = g;
    a.items = b->items_rho; a.nitems = b->nitems_rho; a.queue = b->queue;
    a.W = b->W; a.ek = do_ek ? b->ek : nullptr; a.nktot = b->nktot; a.do_stress = do_stress; a.act = act;
    if (lo) {
        if (mode == RM_Q) rho_t<float, RM_Q>(b, a);
        else if (mode == RM_MU) rho_t<float, RM_MU>(b, a);
        else rho_t<float, RM_QMU>(b, a);
    } else {
        if (mode == RM_Q) rho_t<double, RM_Q>(b, a);
        else if (mode == RM_MU) rho_t<double, RM_MU>(b, a);
        else if (mode == RM_QMU) rho_t<double, RM_QMU>(b, a);
        else rho_t<double, RM_ADJ>(b, a);
    }
}

template <typename real, int NCH>
static void back_t(mdff_batch *b, BackBArgs &a)
{
    int grid = std::max(1, std::min(a.nitems, b->bpc * b->ncu));
    switch (b->ipb) {
    case 128: hipLaunchKernelGGL((kb_back<real, NCH, 128>), dim3(grid), dim3(256), 0, b->st, a); break;
    case 256: hipLaunchKernelGGL((kb_back<real, NCH, 256>), dim3(grid), dim3(256), 0, b->st, a); break;
    default: hipLaunchKernelGGL((kb_back<real, NCH, 64>), dim3(grid), dim3(256), 0, b->st, a); break;
    }
}

// field (and gradient, force) of W at every ion of the active configurations
static void launch_back(mdff_batch *b, int lo, int nch, const double *q, const double *mu, double *ef, double *efg,
                        double *f, const int *act, bool exact = false)
{
    if (b->mf) {
        // moment-weighted W columns in the precision the sum reads (FP64
        // unless the FP32 matrix instructions run), then the back-projection
        const int CH = nch == 9 ? 6 : 3;
        const int form = !lo ? 0 : (exact ? 1 : 2);
        const dim3 gw(16, b->nconf);
        if (form == 2) hipLaunchKernelGGL(kb_mf_wm<float>, gw, dim3(256), 0, b->st, b->conf, b->boxk, b->W, CH, (float *)b->wm, act);
        else hipLaunchKernelGGL(kb_mf_wm<double>, gw, dim3(256), 0, b->st, b->conf, b->boxk, b->W, CH, b->wm, act);
        MBackArgs m;
        m.conf = b->conf;
        m.tabT = lo ? (const void *)b->tabT32 : (const void *)b->tabT64;
        m.wm = b->wm;
        m.by = b->by; m.bz = b->bz; m.n = b->n; m.CH = CH;
        m.q = q; m.mu = mu; m.ef = ef; m.efg = efg; m.f = f; m.act = act;
        for (size_t gi = 0; gi < b->rng_mb.size();) {
            const int CT = bmf_ct(b->kx_mb[gi], CH);
            size_t gj = gi;
            while (gj + 1 < b->rng_mb.size() && bmf_ct(b->kx_mb[gj + 1], CH) == CT) gj++;
            m.items = b->items_mb + b->rng_mb[gi].first;
            m.nitems = b->rng_mb[gj].second - b->rng_mb[gi].first;
            const dim3 gb(bdiv(m.nitems, 4));
#define MBACK(TT, WT, RT, C_, R_) hipLaunchKernelGGL((kb_mf_back<TT, WT, RT, C_, R_>), gb, dim3(256), 0, b->st, m)
#define MBACK_R(TT, WT, RT, C_) do { if (b->mfr == 2) MBACK(TT, WT, RT, C_, 2); else MBACK(TT, WT, RT, C_, 1); } while (0)
#define MBACK_T(TT, WT, RT) do { if (CT <= 2) MBACK_R(TT, WT, RT, 2); else if (CT <= 4) MBACK_R(TT, WT, RT, 4); \
                                 else if (CT <= 6) MBACK_R(TT, WT, RT, 6); else if (CT <= 8) MBACK_R(TT, WT, RT, 8); \
                                 else if (CT <= 10) MBACK_R(TT, WT, RT, 10); else MBACK_R(TT, WT, RT, 12); } while (0)
            if (form == 0) MBACK_T(double, double, double);
            else if (form == 1) MBACK_T(float, double, double);
            else MBACK_T(float, float, float);
#undef MBACK_T
#undef MBACK_R
#undef MBACK
            gi = gj + 1;
        }
        return;
    }
    BackBArgs a;
    a.conf = b->conf; a.kv = b->kv;
    a.tab = lo ? (const void *)b->tabT32 : (const void *)b->tabT64;
    a.TS = b->TS; a.by = b->by; a.bz = b->bz; a.n = b->n;
    a.W = b->W; a.items = b->items_back; a.nitems = b->nitems_back; a.queue = b->queue + 2;
    a.q = q; a.mu = mu; a.ef = ef; a.efg = efg; a.f = f; a.act = act;
    if (lo) { if (nch == 9) back_t<float, 9>(b, a); else back_t<float, 3>(b, a); }
    else { if (nch == 9) back_t<double, 9>(b, a); else back_t<double, 3>(b, a); }
}

static DirArgs bdir_args(mdff_batch *b, const double *q, const double *mu, int lcouldamp, int q_zero, const int *act)
{
    DirArgs a;
    memset(&a, 0, sizeof(a));
    a.n = b->n; a.cap = b->cap; a.s = b->s; a.type = b->type; a.nbr = b->nbr; a.cnt = b->cnt; a.T = b->tab_d;
    a.q = q; a.mu = mu; a.lcouldamp = lcouldamp; a.active = act; a.q_zero = q_zero;
    a.ef = b->ef_dir; a.efg = b->efg_dummy; a.f = b->f_dir; a.uw = b->uw_dir;
    a.conf = b->conf; a.cfg = b->cfg;
    return a;
}

template <typename real, bool EF, bool F>
static void bdir_l(mdff_batch *b, const DirArgs &a)
{
    unsigned blocks = bdiv((size_t)b->n * b->lanes, 256);
    if (b->lanes == 32)
        hipLaunchKernelGGL((k_ewald_dir<real, 32, EF, false, false, F, true>), dim3(blocks), dim3(256), 0, b->st, a);
    else if (b->lanes == 64)
        hipLaunchKernelGGL((k_ewald_dir<real, 64, EF, false, false, F, true>), dim3(blocks), dim3(256), 0, b->st, a);
    else
        hipLaunchKernelGGL((k_ewald_dir<real, 16, EF, false, false, F, true>), dim3(blocks), dim3(256), 0, b->st, a);
}

// real-space Ewald: field only (ef) or forces + energy/virial (f)
static void launch_bdir(mdff_batch *b, const DirArgs &a, int lo, int forces)
{
    if (lo) { if (forces) bdir_l<float, false, true>(b, a); else bdir_l<float, true, false>(b, a); }
    else { if (forces) bdir_l<double, false, true>(b, a); else bdir_l<double, true, false>(b, a); }
}

static void launch_bshort(mdff_batch *b)
{
    ShortArgs a;
    memset(&a, 0, sizeof(a));
    a.n = b->n; a.cap = b->cap; a.s = b->s; a.type = b->type; a.nbr = b->nbr; a.cnt = b->cnt; a.T = b->tab_d;
    a.f = b->f_sr; a.acc = b->acc_sr; a.conf = b->conf; a.cfg = b->cfg;
    unsigned blocks = bdiv((size_t)b->n * b->lanes, 256);
    bool lo = lo_b(b);
    if (b->lanes == 32) {
        if (lo) hipLaunchKernelGGL((k_short<float, 32, true>), dim3(blocks), dim3(256), 0, b->st, a);
        else hipLaunchKernelGGL((k_short<double, 32, true>), dim3(blocks), dim3(256), 0, b->st, a);
    } else if (b->lanes == 64) {
        if (lo) hipLaunchKernelGGL((k_short<float, 64, true>), dim3(blocks), dim3(256), 0, b->st, a);
        else hipLaunchKernelGGL((k_short<double, 64, true>), dim3(blocks), dim3(256), 0, b->st, a);
    } else {
        if (lo) hipLaunchKernelGGL((k_short<float, 16, true>), dim3(blocks), dim3(256), 0, b->st, a);
        else hipLaunchKernelGGL((k_short<double, 16, true>), dim3(blocks), dim3(256), 0, b->st, a);
    }
}

// ------------------------------------------------------------ phases
static int build_list(mdff_batch *b)
{
    int n = b->n;
    hipLaunchKernelGGL(kb_frac, dim3(bdiv(n, 256)), dim3(256), 0, b->st, n, b->conf, b->cfg, b->r, b->s);
    for (int attempt = 0; attempt < 3; attempt++) {
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, b->st, b->ctl, 0);
        hipLaunchKernelGGL(kb_list, dim3(bdiv(n, 4)), dim3(256), 0, b->st, n, b->conf, b->cfg, b->s, b->cap, b->nbr,
                           b->cnt, b->ctl);
        int ov = 0;
        BCHK(hipMemcpyAsync(b->any_h, b->ctl, sizeof(int), hipMemcpyDeviceToHost, b->st));
        BCHK(hipStreamSynchronize(b->st));
        ov = *b->any_h;
        if (!ov) return MDFF_OK;
        bfree(b->nbr);
        b->cap = (ov + ov / 4 + 63) / 64 * 64;
        // no zero fill: the build writes every entry below each ion's count
        if (hipMalloc((void **)&b->nbr, sizeof(unsigned) * (size_t)n * b->cap) != hipSuccess) {
            b->nbr = nullptr;
            return MDFF_ERR_NOMEM;
        }
    }
    return MDFF_ERR_LIST_OVERFLOW;
}

// field of the dipoles mu at every ion of the active configurations, plus
// base (exact: FP64 k-space arithmetic in the mixed form)
static void induced_field(mdff_batch *b, int lo, const double *mu, const double *base, double *ef, const int *act,
                          bool exact = true)
{
    int n = b->n;
    DirArgs a = bdir_args(b, b->qz, mu, 0, 1, act);
    launch_bdir(b, a, lo, 0);
    launch_rho(b, lo, RM_MU, b->qz, mu, nullptr, 0, 0, act, exact);
    launch_back(b, lo, 3, nullptr, nullptr, b->ef_rec, nullptr, nullptr, act, exact);
    hipLaunchKernelGGL(kb_field_sum, dim3(bdiv(3 * n, 256)), dim3(256), 0, b->st, n, b->cfg, act, b->conf, mu,
                       b->ef_dir, b->ef_rec, base, ef);
}

static int read_any(mdff_batch *b)
{
    BCHK(hipMemcpyAsync(b->any_h, b->any, sizeof(int), hipMemcpyDeviceToHost, b->st));
    BCHK(hipStreamSynchronize(b->st));
    return MDFF_OK;
}

// momentpolaSCF pim.c:177-306 for every configuration at once; each
// configuration follows its own iteration sequence and stops on its own test
static int scf(mdff_batch *b)
{
    int n = b->n, nc = b->nconf;
    int maxit = b->p.max_scf_pim_iter, minit = b->p.min_scf_pim_iter;
    hipLaunchKernelGGL(kb_set_conf_int, dim3(bdiv(nc, 256)), dim3(256), 0, b->st, nc, b->act,
                       (maxit > 0 || minit > 0) ? 1 : 0);
    hipLaunchKernelGGL(kb_set_conf_int, dim3(bdiv(nc, 256)), dim3(256), 0, b->st, nc, b->iscf, 0);
    const int lo = lo_b(b);
    if (b->cg) {
        // the mixed form: conjugate gradients per configuration (the system
        // and its carried residual: kb_cg_*), T applied in FP32 (FP32 real
        // space and matrix-core k-space) to the search direction only;
        // iteration 0 forms E_0 = E_stat + T mu_0 whole
        const dim3 g1(bdiv(n, 256)), g3(bdiv(3 * n, 256));
        hipLaunchKernelGGL(kb_scf_mu, g1, dim3(256), 0, b->st, n, b->cfg, b->act, 1, b->type, b->tab_d, b->hist, b->ef,
                           b->mu, b->upol_i);
        induced_field(b, lo, b->mu, b->ef_stat, b->ef, b->act, false);
        hipLaunchKernelGGL(kb_cg_start, g1, dim3(256), 0, b->st, n, b->cfg, b->act, b->type, b->tab_d, b->ef, b->mu,
                           b->cg_r, b->cg_z, b->cg_p, b->cg_part, 0);
        segsum(b, b->cg_part, n, 2, 0, b->cg_s);
        hipLaunchKernelGGL(kb_cg_control, dim3(1), dim3(1024), 0, b->st, nc, b->conf, b->cg_s, 1, b->act, b->iscf,
                           b->rmsd, b->cg_rz, b->cg_beta, maxit, minit, b->p.conv_tol_ind, b->any,
                           (const double *)nullptr);
        b->sweeps++;
        int done = 1;
        const int lim = std::max(maxit, minit) + 1;
        while (done < lim) {
            int rc = read_any(b);
            if (rc) return rc;
            if (!*b->any_h) break;
            for (int it = 0; it < 4 && done < lim; it++, done++) {
                induced_field(b, lo, b->cg_p, nullptr, b->cg_w, b->act, false);
                hipLaunchKernelGGL(kb_cg_pq, g1, dim3(256), 0, b->st, n, b->cfg, b->act, b->type, b->tab_d, b->cg_p,
                                   b->cg_w, b->cg_q, b->cg_part);
                segsum(b, b->cg_part, n, 1, 0, b->cg_pq);
                hipLaunchKernelGGL(kb_cg_update, g1, dim3(256), 0, b->st, n, b->cfg, b->act, b->type, b->tab_d,
                                   b->cg_rz, b->cg_pq, b->cg_p, b->cg_q, b->cg_w, b->mu, b->cg_r, b->ef, b->cg_z,
                                   b->cg_part, 0);
                segsum(b, b->cg_part, n, 2, 0, b->cg_s);
                hipLaunchKernelGGL(kb_cg_control, dim3(1), dim3(1024), 0, b->st, nc, b->conf, b->cg_s, 0, b->act,
                                   b->iscf, b->rmsd, b->cg_rz, b->cg_beta, maxit, minit, b->p.conv_tol_ind, b->any,
                                   (const double *)nullptr);
                hipLaunchKernelGGL(kb_cg_dir, g3, dim3(256), 0, b->st, n, b->cfg, b->act, b->cg_beta, b->cg_z, b->cg_p);
                b->sweeps++;
            }
        }
        hipLaunchKernelGGL(kb_cg_upol, g1, dim3(256), 0, b->st, n, b->type, b->tab_d, b->mu, b->upol_i);
        BCHK(hipMemcpyAsync(b->hist, b->mu, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, b->st));
        return MDFF_OK;
    }
    // the FP64 form (and a mixed batch whose polarisabilities are not
    // symmetric) follows the reference's Jacobi sequence; the mixed form's
    // fields then keep FP64 k-space arithmetic on its FP32 tables (linear in
    // mu to FP64 rounding, as the stop test at conv_tol_ind needs)
    int limit = std::max(maxit, minit) + 1;
    int launched = 0;
    while (launched < limit) {
        int chunk = launched == 0 ? std::max(1, std::min(minit, 4)) : 4;
        for (int it = 0; it < chunk && launched < limit; it++, launched++) {
            hipLaunchKernelGGL(kb_scf_mu, dim3(bdiv(n, 256)), dim3(256), 0, b->st, n, b->cfg, b->act,
                               launched == 0 ? 1 : 0, b->type, b->tab_d, b->hist, b->ef, b->mu, b->upol_i);
            induced_field(b, lo, b->mu, b->ef_stat, b->ef, b->act);
            hipLaunchKernelGGL(kb_scf_rmsd, dim3(bdiv(n, 256)), dim3(256), 0, b->st, n, b->cfg, b->act, b->type,
                               b->tab_d, b->mu, b->ef, b->rmsd_i);
            segsum(b, b->rmsd_i, n, 1, 0, b->rmsd2);
            hipLaunchKernelGGL(kb_scf_control, dim3(1), dim3(1024), 0, b->st, nc, b->conf, b->rmsd2, b->act,
                               b->iscf, b->rmsd, maxit, minit, b->p.conv_tol_ind, b->any);
            b->sweeps++;
        }
        int rc = read_any(b);
        if (rc) return rc;
        if (!*b->any_h) break;
    }
    // dipia_ind[0] = mu (pim.c:281-285): the next call starts here
    BCHK(hipMemcpyAsync(b->hist, b->mu, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, b->st));
    return MDFF_OK;
}

// one static engforce per configuration.  Per-configuration sums, host
// side seg_h[c*SG_N + col]: [0..13] short range (u_bh, w_bh[6], u_lj,
// w_lj[6]), [14..20] dir energy/virial, [21..27] rec, [28..29] self, [30] upol.
// Device side, by region of seg: short range nc x 14 at 0, upol nc at 14 nc,
// dir nc x 7 at 16 nc, rec nc x 7 at 24 nc, self nc x 2 at 32 nc.
enum { SG_SR = 0, SG_DIR = 14, SG_REC = 21, SG_SELF = 28, SG_UPOL = 30, SG_N = 32 };

static int engforce_b(mdff_batch *b)
{
    int n = b->n, nc = b->nconf;
    int lo = lo_b(b);
    BCHK(hipEventRecord(b->ev[0], b->st));
    int rc = build_list(b);
    if (rc) return rc;
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, b->st, b->ctl + C_NAN, 0);
    hipLaunchKernelGGL(k_zero, dim3(bdiv((size_t)nc * 40, 256)), dim3(256), 0, b->st, b->seg, (size_t)nc * 40);
    hipLaunchKernelGGL(k_zero, dim3(bdiv(3 * (size_t)n, 256)), dim3(256), 0, b->st, b->f_sr, (size_t)3 * n);
    hipLaunchKernelGGL(k_zero, dim3(bdiv(14 * (size_t)n, 256)), dim3(256), 0, b->st, b->acc_sr, (size_t)14 * n);
    if (b->lnonbonded) launch_bshort(b);
    segsum(b, b->acc_sr, n, 14, 0, b->seg);
    hipLaunchKernelGGL(k_zero, dim3(bdiv(3 * (size_t)n, 256)), dim3(256), 0, b->st, b->f_dir, (size_t)3 * n);
    hipLaunchKernelGGL(k_zero, dim3(bdiv(3 * (size_t)n, 256)), dim3(256), 0, b->st, b->f_rec, (size_t)3 * n);
    for (auto &v : b->nscf_h) v = 0;
    for (auto &v : b->rmsd_h) v = 0.0;
    std::vector<double> sr((size_t)nc * 14), sd((size_t)nc * 7), sk((size_t)nc * 7), ss((size_t)nc * 2),
        su((size_t)nc, 0.0);
    BCHK(hipMemcpyAsync(sr.data(), b->seg, sizeof(double) * nc * 14, hipMemcpyDeviceToHost, b->st));
    if (b->p.lcoulombic) {
        if (lo) hipLaunchKernelGGL(kb_eikr<float>, dim3(bdiv(n, 128)), dim3(128), 0, b->st, n, b->conf, b->cfg, b->r,
                                   b->TS, b->by, b->bz, b->tab32, b->tabT32);
        hipLaunchKernelGGL(kb_eikr<double>, dim3(bdiv(n, 128)), dim3(128), 0, b->st, n, b->conf, b->cfg, b->r, b->TS,
                           b->by, b->bz, b->tab64, b->tabT64);
        if (b->lpol) {
            // static field of the charges (momentpolaSCF pim.c:216-218), damped
            DirArgs a = bdir_args(b, b->q, b->mu0, b->ldmp, 0, nullptr);
            launch_bdir(b, a, lo, 0);
            launch_rho(b, lo, RM_Q, b->q, nullptr, nullptr, 0, 0, nullptr);
            launch_back(b, lo, 3, nullptr, nullptr, b->ef_rec, nullptr, nullptr, nullptr);
            hipLaunchKernelGGL(kb_field_sum, dim3(bdiv(3 * n, 256)), dim3(256), 0, b->st, n, b->cfg, nullptr, b->conf,
                               b->mu0, b->ef_dir, b->ef_rec, nullptr, b->ef_stat);
            rc = scf(b);
            if (rc) return rc;
            BCHK(hipMemcpyAsync(b->nscf_h.data(), b->iscf, sizeof(int) * nc, hipMemcpyDeviceToHost, b->st));
            BCHK(hipMemcpyAsync(b->rmsd_h.data(), b->rmsd, sizeof(double) * nc, hipMemcpyDeviceToHost, b->st));
            segsum(b, b->upol_i, n, 1, 0, b->seg + (size_t)nc * 14);
            BCHK(hipMemcpyAsync(su.data(), b->seg + (size_t)nc * 14, sizeof(double) * nc, hipMemcpyDeviceToHost, b->st));
        } else {
            hipLaunchKernelGGL(k_zero, dim3(bdiv(3 * (size_t)n, 256)), dim3(256), 0, b->st, b->mu, (size_t)3 * n);
        }
        // final multipole_ES: forces, energy, stress (field.c:381-383)
        DirArgs a = bdir_args(b, b->q, b->mu, b->ldmp, 0, nullptr);
        launch_bdir(b, a, lo, 1);
        segsum(b, b->uw_dir, n, 7, 0, b->seg + (size_t)nc * 16);
        int mode = b->lpol ? RM_QMU : RM_Q;
        launch_rho(b, lo, mode, b->q, b->mu, nullptr, 1, 1, nullptr);
        segsum(b, b->ek, b->nktot, 7, 1, b->seg + (size_t)nc * 24);
        launch_back(b, lo, b->lpol ? 9 : 3, b->q, b->mu, b->ef_rec, b->lpol ? b->efg_rec : nullptr, b->f_rec, nullptr);
        hipLaunchKernelGGL(k_self_terms, dim3(bdiv(n, 256)), dim3(256), 0, b->st, n, b->q, b->mu, b->selfb);
        segsum(b, b->selfb, n, 2, 0, b->seg + (size_t)nc * 32);
        BCHK(hipMemcpyAsync(sd.data(), b->seg + (size_t)nc * 16, sizeof(double) * nc * 7, hipMemcpyDeviceToHost, b->st));
        BCHK(hipMemcpyAsync(sk.data(), b->seg + (size_t)nc * 24, sizeof(double) * nc * 7, hipMemcpyDeviceToHost, b->st));
        BCHK(hipMemcpyAsync(ss.data(), b->seg + (size_t)nc * 32, sizeof(double) * nc * 2, hipMemcpyDeviceToHost, b->st));
    }
    hipLaunchKernelGGL(k_assemble_force, dim3(bdiv(3 * (size_t)n, 256)), dim3(256), 0, b->st, 3 * n, b->f_sr,
                       b->f_dir, b->f_rec, b->p.lcoulombic ? b->u.coul_unit : 0.0, b->f, b->ctl);
    BCHK(hipEventRecord(b->ev[1], b->st));
    BCHK(hipStreamSynchronize(b->st));
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, b->ev[0], b->ev[1]) == hipSuccess) b->last_ms = ms;
    // per-configuration observables (thermo.c:12-33 for a static call)
    for (int c = 0; c < nc; c++) {
        double *o = b->seg_h + (size_t)c * SG_N;
        memset(o, 0, sizeof(double) * SG_N);
        for (int k = 0; k < 14; k++) o[SG_SR + k] = sr[(size_t)c * 14 + k];
        for (int k = 0; k < 7; k++) { o[SG_DIR + k] = sd[(size_t)c * 7 + k]; o[SG_REC + k] = sk[(size_t)c * 7 + k]; }
        o[SG_SELF] = ss[(size_t)c * 2];
        o[SG_SELF + 1] = ss[(size_t)c * 2 + 1];
        o[SG_UPOL] = su[c];
    }
    return MDFF_OK;
}

static void thermo_of(mdff_batch *b, int c, mdff_thermo *th)
{
    const double *S = b->seg_h + (size_t)c * SG_N;
    const BConfD &K = b->conf_h[c];
    mdff_thermo t;
    memset(&t, 0, sizeof(t));
    t.rmsd_fp64 = -1.0;
    double iv = K.iv;
    auto sym = [&](const double *w, double scale, double out[3][3]) {
        out[0][0] = w[0] * scale; out[0][1] = w[1] * scale; out[0][2] = w[2] * scale;
        out[1][0] = w[1] * scale; out[1][1] = w[3] * scale; out[1][2] = w[4] * scale;
        out[2][0] = w[2] * scale; out[2][1] = w[4] * scale; out[2][2] = w[5] * scale;
    };
    if (b->p.lbhmft || b->p.lbhmftd) {
        t.u_bhmftd = S[0];
        sym(S + 1, iv, t.tau_bhmftd);
        t.pvir_bhmftd = (t.tau_bhmftd[0][0] + t.tau_bhmftd[1][1] + t.tau_bhmftd[2][2]) / 3.0;
    }
    if (b->p.lnmlj) {
        t.u_lj = S[7];
        sym(S + 8, iv, t.tau_lj);
        t.pvir_lj = (t.tau_lj[0][0] + t.tau_lj[1][1] + t.tau_lj[2][2]) / 3.0;
    }
    if (b->p.lcoulombic) {
        double cu = b->u.coul_unit;
        double u_self = -K.selfa * S[SG_SELF] - K.selfa2 * S[SG_SELF + 1];
        t.u_coul = (S[SG_DIR] + S[SG_REC] * K.ttpiV + u_self) * cu;
        double td[3][3], tr[3][3];
        sym(S + SG_DIR + 1, 0.5 * iv, td);
        sym(S + SG_REC + 1, K.ttpiV * iv, tr);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) t.tau_coul[i][j] = (td[i][j] + tr[i][j]) * cu;
        t.pvir_coul = (t.tau_coul[0][0] + t.tau_coul[1][1] + t.tau_coul[2][2]) / 3.0;
        t.u_pol = b->lpol ? S[SG_UPOL] * cu : 0.0;
    }
    t.u_tot = t.u_lj + t.u_bhmftd + t.u_coul + t.u_pol;
    t.e_tot = t.u_tot;
    t.h_tot = t.u_tot;
    t.pressure = t.pvir_lj + t.pvir_bhmftd + t.pvir_coul;
    t.n_scf = b->nscf_h[c];
    t.rmsd = b->rmsd_h[c];
    *th = t;
}

// ------------------------------------------------------------ C ABI
void mdff_batch_destroy(mdff_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->st) (void)hipStreamSynchronize(b->st);
    BConfD *c0 = b->conf;
    bfree(c0); b->conf = nullptr;
    bfree(b->tab_d); bfree(b->cfg); bfree(b->type); bfree(b->r); bfree(b->s); bfree(b->q); bfree(b->qz);
    bfree(b->mu); bfree(b->mu0); bfree(b->hist); bfree(b->ef_stat); bfree(b->ef); bfree(b->ef_dir); bfree(b->ef_rec);
    bfree(b->efg_rec); bfree(b->efg_dummy); bfree(b->f); bfree(b->f_sr); bfree(b->f_dir); bfree(b->f_rec);
    bfree(b->acc_sr); bfree(b->uw_dir); bfree(b->selfb); bfree(b->upol_i); bfree(b->rmsd_i);
    bfree(b->nbr); bfree(b->cnt); bfree(b->ctl);
    bfree(b->kv); bfree(b->W); bfree(b->ek); bfree(b->tab64); bfree(b->tab32); bfree(b->tabT64); bfree(b->tabT32); bfree(b->queue);
    bfree(b->items_rho); bfree(b->items_back); bfree(b->act); bfree(b->iscf); bfree(b->any);
    bfree(b->rmsd2); bfree(b->rmsd); bfree(b->seg); bfree(b->g); bfree(b->sdir); bfree(b->srhs); bfree(b->lam);
    bfree(b->eT); bfree(b->edel); bfree(b->cols); bfree(b->tsum);
    bfree(b->boxk); bfree(b->kcfg); bfree(b->Sk); bfree(b->items_mf); bfree(b->items_mb);
    bfree(b->aw); bfree(b->wm);
    bfree(b->cg_r); bfree(b->cg_z); bfree(b->cg_p); bfree(b->cg_w); bfree(b->cg_q); bfree(b->cg_part);
    bfree(b->cg_s); bfree(b->cg_pq); bfree(b->cg_rz); bfree(b->cg_beta); bfree(b->cg_e);
    if (b->any_h) (void)hipHostFree(b->any_h);
    if (b->seg_h) free(b->seg_h);
    for (int i = 0; i < 4; i++)
        if (b->ev[i]) (void)hipEventDestroy(b->ev[i]);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
}

int mdff_batch_create(mdff_batch **out, const mdff_params *p, int nconf, const int *nion, const int *typia,
                      const double *A9, const double *rx, const double *ry, const double *rz,
                      const mdff_batch_ewald *ew, int precision, int device)
{
    if (!out || !p || nconf < 1 || !nion || !typia || !A9 || !rx || !ry || !rz) return MDFF_ERR_ARG;
    *out = nullptr;
    if (precision != MDFF_FP64 && precision != MDFF_MIXED) return MDFF_ERR_ARG;
    if (p->ntype < 1 || p->ntype > NT) return MDFF_ERR_CONFIG;
    for (int t = 0; t < p->ntype; t++)
        for (int k = 0; k < 3; k++)
            if (p->dipit[t][k] != 0.0) return MDFF_ERR_UNSUPPORTED;
    for (int t = 0; t < p->ntype; t++)
        if (p->lcoulombic && p->lpolar[t] && p->algo_pim != MDFF_PIM_SCF) return MDFF_ERR_UNSUPPORTED;
    long long ntot = 0;
    for (int c = 0; c < nconf; c++) {
        if (nion[c] < 1) return MDFF_ERR_ARG;
        ntot += nion[c];
    }
    if (ntot > (long long)NBR_IDX) return MDFF_ERR_ARG;
    int n = (int)ntot;
    for (int i = 0; i < n; i++)
        if (typia[i] < 0 || typia[i] >= p->ntype) return MDFF_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MDFF_ERR_DEVICE;
    if (device < 0 || device >= ndev) return MDFF_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return MDFF_ERR_DEVICE;

    mdff_batch *b = new mdff_batch();
    b->p = *p;
    b->nconf = nconf;
    b->n = n;
    b->ntype = p->ntype;
    b->precision = precision;
    b->device = device;
    mdff_units_get(p, &b->u);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        b->ncu = prop.multiProcessorCount;
    for (int t = 0; t < b->ntype; t++) {
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++)
                if (p->polit[t][j][k] != 0.0) b->lpol = 1;
        for (int i = 0; i < b->ntype; i++)
            for (int k = 0; k < b->ntype; k++)
                if (p->lpoldamping[t][i][k]) b->ldmp = 1;
    }
    if (!p->lcoulombic) b->lpol = 0;
    b->lnonbonded = p->lnmlj || p->lbhmft || p->lbhmftd;
    typetab_build(*p, b->ntype, p->cutshortrange * p->cutshortrange, b->tab);
    b->typ_h.assign(typia, typia + n);
    // configurations
    b->conf_h.resize(nconf);
    std::vector<KVec> kvh;
    int off = 0, Kxm = 0, Kym = 0, Kzm = 0;
    double rcmax = 0.0, rhomax = 0.0;
    for (int c = 0; c < nconf; c++) {
        BConfD &K = b->conf_h[c];
        memset(&K, 0, sizeof(K));
        double A[3][3];
        for (int i = 0; i < 9; i++) A[i / 3][i % 3] = A9[(size_t)9 * c + i];
        mdff_cell cell;
        mdff_lattice(A, b->u.press_unit, &cell);
        memcpy(K.c.A, cell.A, sizeof(K.c.A));
        memcpy(K.c.B, cell.B, sizeof(K.c.B));
        double alpha = p->alphaES, rl = p->cutlongrange, rs = p->cutshortrange;
        int kes[3] = {p->kES[0], p->kES[1], p->kES[2]};
        if (ew) {
            alpha = ew[c].alpha; rl = ew[c].rcut_long; rs = ew[c].rcut_short;
            for (int k = 0; k < 3; k++) kes[k] = ew[c].kES[k];
        } else if (p->lcoulombic && p->lautoES) {
            mdff_auto_ewald(p, &cell, &alpha, kes);
        }
        double wmin = std::min(cell.w[0], std::min(cell.w[1], cell.w[2]));
        if (rl > 0.5 * wmin * (1.0 + 1e-12) || rs > 0.5 * wmin * (1.0 + 1e-12) || kes[0] < 0 || kes[1] < 0 ||
            kes[2] < 0 || (p->lcoulombic && alpha <= 0.0)) {
            delete b;
            return MDFF_ERR_CONFIG;
        }
        K.alpha = alpha;
        K.rc2s = rs * rs;
        K.rc2l = rl * rl;
        K.selfa = alpha / sqrt(MDFF_PI);
        K.selfa2 = 2.0 * K.selfa * alpha * alpha / 3.0;
        K.ttpiV = 2.0 * MDFF_TPI * cell.inveOmega;
        K.tfpiV = 2.0 * K.ttpiV;
        K.iv = cell.inveOmegaPU;
        K.off = off;
        K.n = nion[c];
        int np = 0;
        for (int i = off; i < off + nion[c]; i++)
            if (b->lpol && p->lpolar[typia[i]]) np++;
        K.npol = np;
        K.Kx = kes[0]; K.Ky = kes[1]; K.Kz = kes[2];
        K.koff = (int)kvh.size();
        K.boxoff = (int)b->nbox;
        K.awoff = (int)b->naw;
        K.wmoff = (int)b->nwm;
        if (p->lcoulombic) {
            double a2 = alpha * alpha;
            const size_t box = (size_t)(K.Kx + 1) * (2 * K.Ky + 1) * (2 * K.Kz + 1);
            b->boxk_h.resize(b->nbox + box, -1);
            b->nbox += box;
            // operand blocks of the matrix-core sums, sized for 4 channels
            // and 6 moments (the largest tile counts)
            const size_t Pbox = (size_t)(2 * K.Ky + 1) * (2 * K.Kz + 1);
            b->naw += (size_t)2 * nion[c] * 16 * bmf_mt(K.Kx, 4);
            b->nwm += (size_t)2 * ((Pbox + 1) & ~(size_t)1) * 16 * bmf_ct(K.Kx, 6);
            size_t e = K.boxoff;
            for (int nx = 0; nx <= K.Kx; nx++)
                for (int ny = -K.Ky; ny <= K.Ky; ny++)
                    for (int nz = -K.Kz; nz <= K.Kz; nz++, e++) {
                        bool in_half = (nx > 0) || (nx == 0 && ny > 0) || (nx == 0 && ny == 0 && nz > 0);
                        if (!in_half) continue;
                        const double(*B)[3] = cell.B;
                        KVec v;
                        v.kx = MDFF_TPI * (nx * B[0][0] + ny * B[0][1] + nz * B[0][2]);
                        v.ky = MDFF_TPI * (nx * B[1][0] + ny * B[1][1] + nz * B[1][2]);
                        v.kz = MDFF_TPI * (nx * B[2][0] + ny * B[2][1] + nz * B[2][2]);
                        double kk = v.kx * v.kx + v.ky * v.ky + v.kz * v.kz;
                        if (kk == 0.0) continue;
                        v.Ak = exp(-kk * 0.25 / a2) / kk;
                        v.kcoe = 2.0 * (1.0 / kk + 1.0 / a2 / 4.0);
                        v.nx = nx; v.ny = ny; v.nz = nz; v.pad = 0;
                        b->boxk_h[e] = (int)kvh.size() - K.koff;
                        kvh.push_back(v);
                    }
        }
        K.nk = (int)kvh.size() - K.koff;
        Kxm = std::max(Kxm, K.Kx); Kym = std::max(Kym, K.Ky); Kzm = std::max(Kzm, K.Kz);
        rcmax = std::max(rcmax, std::max(rl, rs));
        rhomax = std::max(rhomax, (double)nion[c] * cell.inveOmega);
        off += nion[c];
    }
    b->nktot = kvh.size();
    b->TS = (Kxm + 1) + (2 * Kym + 1) + (2 * Kzm + 1);
    b->by = (Kxm + 1) + Kym;
    b->bz = (Kxm + 1) + (2 * Kym + 1) + Kzm;
    b->nscf_h.assign(nconf, 0);
    b->rmsd_h.assign(nconf, 0.0);
    // neighbour capacity from the densest configuration
    double est = 4.0 / 3.0 * MDFF_PI * rcmax * rcmax * rcmax * rhomax;
    b->cap = std::max(64, (int)((est * 1.3 + 32 + 63) / 64) * 64);
    int maxn = 0;
    for (int c = 0; c < nconf; c++) maxn = std::max(maxn, nion[c]);
    b->cap = std::min(b->cap, (maxn + 63) / 64 * 64);

    int rc = MDFF_OK;
#define BALLOC(ptr, cnt) do { if ((rc = balloc(&(ptr), (cnt))) != MDFF_OK) { mdff_batch_destroy(b); return rc; } } while (0)
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) { delete b; return MDFF_ERR_DEVICE; }
    for (int i = 0; i < 4; i++)
        if (hipEventCreate(&b->ev[i]) != hipSuccess) { mdff_batch_destroy(b); return MDFF_ERR_DEVICE; }
    size_t n3 = (size_t)3 * n;
    BALLOC(b->conf, nconf); BALLOC(b->tab_d, 1); BALLOC(b->cfg, n); BALLOC(b->type, n);
    BALLOC(b->r, n3); BALLOC(b->s, n3); BALLOC(b->q, n); BALLOC(b->qz, n); BALLOC(b->mu, n3); BALLOC(b->mu0, n3);
    BALLOC(b->hist, n3); BALLOC(b->ef_stat, n3); BALLOC(b->ef, n3); BALLOC(b->ef_dir, n3); BALLOC(b->ef_rec, n3);
    BALLOC(b->efg_rec, 6 * (size_t)n); BALLOC(b->efg_dummy, 9 * (size_t)n);
    BALLOC(b->f, n3); BALLOC(b->f_sr, n3); BALLOC(b->f_dir, n3); BALLOC(b->f_rec, n3);
    BALLOC(b->acc_sr, 14 * (size_t)n); BALLOC(b->uw_dir, 7 * (size_t)n); BALLOC(b->selfb, 2 * (size_t)n);
    BALLOC(b->upol_i, n); BALLOC(b->rmsd_i, n);
    BALLOC(b->nbr, (size_t)n * b->cap); BALLOC(b->cnt, n); BALLOC(b->ctl, C_COUNT);
    BALLOC(b->kv, std::max<size_t>(1, b->nktot)); BALLOC(b->W, 2 * std::max<size_t>(1, b->nktot));
    BALLOC(b->ek, 7 * std::max<size_t>(1, b->nktot));
    // matrix-core k-space when every box fits the kernels' tiles:
    // (Kx + 1) 4 2 <= 16 BMF_MT rows and (Kx + 1) 6 2 <= 16 BMF_CT columns
    b->mf = p->lcoulombic && b->nktot > 0 && b->nbox > 0;
    for (int c = 0; c < nconf && b->mf; c++)
        if (b->conf_h[c].Kx + 1 > 16) b->mf = 0;
    if (b->naw > (size_t)INT32_MAX || b->nwm > (size_t)INT32_MAX) b->mf = 0;
    if (b->mf) {
        BALLOC(b->boxk, b->nbox); BALLOC(b->kcfg, b->nktot); BALLOC(b->Sk, 8 * b->nktot);
        BALLOC(b->aw, b->naw); BALLOC(b->wm, b->nwm);
    }
    // conjugate-gradient buffers (used while the polarisabilities allow it,
    // cg_flags)
    cg_flags(b);
    if (b->lpol) {
        BALLOC(b->cg_r, n3); BALLOC(b->cg_z, n3); BALLOC(b->cg_p, n3); BALLOC(b->cg_w, n3); BALLOC(b->cg_q, n3);
        BALLOC(b->cg_e, n3);
        BALLOC(b->cg_part, 2 * (size_t)n); BALLOC(b->cg_s, 2 * (size_t)nconf); BALLOC(b->cg_pq, nconf);
        BALLOC(b->cg_rz, nconf); BALLOC(b->cg_beta, nconf);
    }
    BALLOC(b->tab64, 2 * (size_t)n * b->TS);
    BALLOC(b->tabT64, 2 * (size_t)n * b->TS);
    if (precision == MDFF_MIXED) {
        BALLOC(b->tab32, 2 * (size_t)n * b->TS);
        BALLOC(b->tabT32, 2 * (size_t)n * b->TS);
    }
    BALLOC(b->queue, 4); BALLOC(b->act, nconf); BALLOC(b->iscf, nconf); BALLOC(b->any, 1);
    BALLOC(b->rmsd2, nconf); BALLOC(b->rmsd, nconf); BALLOC(b->seg, (size_t)nconf * 40);
    BALLOC(b->g, n3); BALLOC(b->sdir, n3); BALLOC(b->srhs, n3); BALLOC(b->lam, n3); BALLOC(b->eT, n3);
    BALLOC(b->edel, n3); BALLOC(b->cols, (size_t)(6 * NT + 9 + 2) * n); BALLOC(b->tsum, (size_t)NT * (6 * NT + 9) + 4 * (size_t)nconf + 8);
#undef BALLOC
    if (hipHostMalloc((void **)&b->any_h, sizeof(int) * 4) != hipSuccess) { mdff_batch_destroy(b); return MDFF_ERR_NOMEM; }
    b->seg_h = (double *)calloc((size_t)nconf * SG_N, sizeof(double));
    if (!b->seg_h) { mdff_batch_destroy(b); return MDFF_ERR_NOMEM; }
    // ions
    std::vector<int> hc(n);
    std::vector<double> hq(n), hr(n3);
    for (int c = 0; c < nconf; c++)
        for (int i = b->conf_h[c].off; i < b->conf_h[c].off + b->conf_h[c].n; i++) hc[i] = c;
    for (int i = 0; i < n; i++) hq[i] = p->lcoulombic ? p->qit[typia[i]] : 0.0;
    memcpy(hr.data(), rx, sizeof(double) * n);
    memcpy(hr.data() + n, ry, sizeof(double) * n);
    memcpy(hr.data() + 2 * n, rz, sizeof(double) * n);
    if (hipMemcpy(b->conf, b->conf_h.data(), sizeof(BConfD) * nconf, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->tab_d, &b->tab, sizeof(TypeTab), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->cfg, hc.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->type, typia, sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->q, hq.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->r, hr.data(), sizeof(double) * n3, hipMemcpyHostToDevice) != hipSuccess ||
        (b->nktot && hipMemcpy(b->kv, kvh.data(), sizeof(KVec) * b->nktot, hipMemcpyHostToDevice) != hipSuccess)) {
        mdff_batch_destroy(b);
        return MDFF_ERR_DEVICE;
    }
    if (b->mf) {
        std::vector<int> kc(b->nktot);
        for (int c = 0; c < nconf; c++)
            for (int k = 0; k < b->conf_h[c].nk; k++) kc[(size_t)b->conf_h[c].koff + k] = c;
        if (hipMemcpy(b->boxk, b->boxk_h.data(), sizeof(int) * b->nbox, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(b->kcfg, kc.data(), sizeof(int) * b->nktot, hipMemcpyHostToDevice) != hipSuccess) {
            mdff_batch_destroy(b);
            return MDFF_ERR_DEVICE;
        }
    }
    if ((rc = build_items(b)) != MDFF_OK) { mdff_batch_destroy(b); return rc; }
    // every zero fill of the set-up (null stream) is done before the batch stream runs
    if (hipDeviceSynchronize() != hipSuccess) { mdff_batch_destroy(b); return MDFF_ERR_DEVICE; }
    *out = b;
    return MDFF_OK;
}

int mdff_batch_set_positions(mdff_batch *b, const double *rx, const double *ry, const double *rz)
{
    if (!b || !rx || !ry || !rz) return MDFF_ERR_ARG;
    int n = b->n;
    BCHK(hipMemcpy(b->r, rx, sizeof(double) * n, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->r + n, ry, sizeof(double) * n, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->r + 2 * n, rz, sizeof(double) * n, hipMemcpyHostToDevice));
    return MDFF_OK;
}

// Conjugate gradients (the mixed form's SCF, the FP64 form's adjoint solve)
// need every polarisability tensor symmetric positive definite (then the
// system pol^-T - T of a converging SCF is too); the SCF's also the
// matrix-core sums.  Otherwise both run their Jacobi sequences.
static void cg_flags(mdff_batch *b)
{
    bool spd = b->lpol;
    for (int t = 0; t < b->ntype && spd; t++) {
        if (!b->tab.polar[t]) continue;
        const double *m = b->tab.pol[t];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                if (fabs(m[3 * i + j] - m[3 * j + i]) > 1e-12 * (fabs(m[3 * i + j]) + fabs(m[3 * j + i]))) spd = false;
        const double m2 = m[0] * m[4] - m[1] * m[3];
        const double m3 = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                          m[2] * (m[3] * m[7] - m[4] * m[6]);
        if (!(m[0] > 0.0 && m2 > 0.0 && m3 > 0.0)) spd = false;
    }
    b->cg_sym = spd ? 1 : 0;
    b->cg = (b->precision == MDFF_MIXED && b->mf && spd) ? 1 : 0;
}

int mdff_batch_set_params(mdff_batch *b, const mdff_params *p)
{
    if (!b || !p) return MDFF_ERR_ARG;
    if (p->ntype != b->ntype || p->lcoulombic != b->p.lcoulombic) return MDFF_ERR_CONFIG;
    for (int t = 0; t < b->ntype; t++) {
        if (p->qit[t] != b->p.qit[t] || (p->lpolar[t] != 0) != (b->p.lpolar[t] != 0)) return MDFF_ERR_CONFIG;
        for (int k = 0; k < 3; k++)
            if (p->dipit[t][k] != 0.0) return MDFF_ERR_UNSUPPORTED;
        // the SCF of mdff_batch_create's check: momentpolaSCF only
        if (p->lcoulombic && p->lpolar[t] && p->algo_pim != MDFF_PIM_SCF) return MDFF_ERR_UNSUPPORTED;
    }
    // the terms that exist follow the new parameters (damping, pair
    // potentials); whether any ion is polarisable may not change (it sizes
    // the SCF's buffers)
    int pol = 0, dmp = 0;
    for (int t = 0; t < b->ntype; t++) {
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++)
                if (p->polit[t][j][k] != 0.0) pol = 1;
        for (int i = 0; i < b->ntype; i++)
            for (int k = 0; k < b->ntype; k++)
                if (p->lpoldamping[t][i][k]) dmp = 1;
    }
    if (!p->lcoulombic) pol = 0;
    if (pol != b->lpol) return MDFF_ERR_CONFIG;
    mdff_params keep = b->p;
    b->p = *p;
    b->ldmp = dmp;
    b->lnonbonded = p->lnmlj || p->lbhmft || p->lbhmftd;
    // Ewald settings and cutoffs stay those of creation
    b->p.alphaES = keep.alphaES;
    memcpy(b->p.kES, keep.kES, sizeof(keep.kES));
    b->p.cutlongrange = keep.cutlongrange;
    b->p.cutshortrange = keep.cutshortrange;
    typetab_build(b->p, b->ntype, b->p.cutshortrange * b->p.cutshortrange, b->tab);
    BCHK(hipMemcpy(b->tab_d, &b->tab, sizeof(TypeTab), hipMemcpyHostToDevice));
    cg_flags(b);
    return MDFF_OK;
}

int mdff_batch_reset_dipoles(mdff_batch *b)
{
    if (!b) return MDFF_ERR_ARG;
    BCHK(hipMemsetAsync(b->hist, 0, sizeof(double) * 3 * (size_t)b->n, b->st));
    BCHK(hipStreamSynchronize(b->st));
    return MDFF_OK;
}

int mdff_batch_engforce(mdff_batch *b, mdff_thermo *th)
{
    if (!b) return MDFF_ERR_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return MDFF_ERR_DEVICE;
    int rc = engforce_b(b);
    if (rc) return rc;
    int worst = MDFF_OK;
    for (int c = 0; c < b->nconf; c++) {
        if (th) thermo_of(b, c, th + c);
        if (b->lpol && b->conf_h[c].npol > 0 && b->rmsd_h[c] > b->p.conv_tol_ind) worst = MDFF_ERR_SCF;
    }
    int nanf = 0;
    BCHK(hipMemcpy(&nanf, b->ctl + C_NAN, sizeof(int), hipMemcpyDeviceToHost));
    if (nanf) return MDFF_ERR_NAN;
    return worst;
}

int mdff_batch_get_forces(mdff_batch *b, double *fx, double *fy, double *fz)
{
    if (!b || !fx || !fy || !fz) return MDFF_ERR_ARG;
    int n = b->n;
    BCHK(hipMemcpy(fx, b->f, sizeof(double) * n, hipMemcpyDeviceToHost));
    BCHK(hipMemcpy(fy, b->f + n, sizeof(double) * n, hipMemcpyDeviceToHost));
    BCHK(hipMemcpy(fz, b->f + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost));
    return MDFF_OK;
}

int mdff_batch_get_dipoles(mdff_batch *b, double *mu)
{
    if (!b || !mu) return MDFF_ERR_ARG;
    int n = b->n;
    std::vector<double> h((size_t)3 * n);
    BCHK(hipMemcpy(h.data(), b->mu, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) mu[3 * (size_t)i + a] = h[(size_t)a * n + i];
    return MDFF_OK;
}

// ------------------------------------------------------------ adjoint
template <int L>
static void adj_launch(mdff_batch *b, int which, AdjDirArgs &a, AdjShortArgs &s)
{
    unsigned blocks = bdiv((size_t)b->n * L, 256);
    if (which == 0) hipLaunchKernelGGL(kb_adj_dir<L>, dim3(blocks), dim3(256), 0, b->st, a);
    else if (which == 1) hipLaunchKernelGGL(kb_adj_damp<L>, dim3(blocks), dim3(256), 0, b->st, a);
    else hipLaunchKernelGGL(kb_adj_short<L>, dim3(blocks), dim3(256), 0, b->st, s);
}

static void adj_kernel(mdff_batch *b, int which, AdjDirArgs &a, AdjShortArgs &s)
{
    if (b->lanes == 32) adj_launch<32>(b, which, a, s);
    else if (b->lanes == 64) adj_launch<64>(b, which, a, s);
    else adj_launch<16>(b, which, a, s);
}

int mdff_batch_adjoint(mdff_batch *b, const double *fref_x, const double *fref_y, const double *fref_z, mdff_grad *gr)
{
    if (!b || !fref_x || !fref_y || !fref_z || !gr) return MDFF_ERR_ARG;
    if (b->precision != MDFF_FP64) return MDFF_ERR_UNSUPPORTED;
    int rc = mdff_batch_engforce(b, nullptr);
    if (rc && rc != MDFF_ERR_SCF) return rc;
    int n = b->n, nc = b->nconf, nt = b->ntype;
    size_t n3 = (size_t)3 * n;
    memset(gr, 0, sizeof(*gr));
    // g = 2 (F - F_ref), loss
    BCHK(hipMemcpy(b->eT, fref_x, sizeof(double) * n, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->eT + n, fref_y, sizeof(double) * n, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->eT + 2 * n, fref_z, sizeof(double) * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(kb_adj_g, dim3(bdiv(n, 256)), dim3(256), 0, b->st, n, b->f, b->eT, b->eT + n, b->eT + 2 * n,
                       b->g, b->rmsd_i);
    double *T = b->tsum;
    hipLaunchKernelGGL(k_colsum, dim3(1), dim3(1024), 0, b->st, b->rmsd_i, n, n, T, (const int *)nullptr);
    double loss = 0.0;
    BCHK(hipMemcpyAsync(&loss, T, sizeof(double), hipMemcpyDeviceToHost, b->st));
    BCHK(hipStreamSynchronize(b->st));
    gr->loss = loss;
    double cu = b->u.coul_unit;
    AdjDirArgs a;
    memset(&a, 0, sizeof(a));
    a.n = n; a.cap = b->cap; a.conf = b->conf; a.cfg = b->cfg; a.s = b->s; a.type = b->type; a.nbr = b->nbr;
    a.cnt = b->cnt; a.T = b->tab_d; a.q = b->q; a.mu = b->mu; a.g = b->g; a.lam = b->lam; a.lcouldamp = b->ldmp;
    a.nt = nt; a.coul = cu;
    AdjShortArgs sa;
    memset(&sa, 0, sizeof(sa));
    sa.n = n; sa.cap = b->cap; sa.nt = nt; sa.conf = b->conf; sa.cfg = b->cfg; sa.s = b->s; sa.type = b->type;
    sa.nbr = b->nbr; sa.cnt = b->cnt; sa.T = b->tab_d; sa.g = b->g;
    const int NCOL = 6 * nt + 9;
    double *cols = b->cols;
    hipLaunchKernelGGL(k_zero, dim3(bdiv((size_t)NCOL * n, 256)), dim3(256), 0, b->st, cols, (size_t)NCOL * n);
    // short-range parameter terms
    if (b->lnonbonded && (b->p.lbhmft || b->p.lbhmftd)) {
        sa.out = cols;
        adj_kernel(b, 2, a, sa);
    }
    int iters_max = 0;
    double resid_max = 0.0;
    if (b->p.lcoulombic && b->lpol) {
        // right-hand side D_g E at the polarisable ions
        a.out = b->sdir;
        adj_kernel(b, 0, a, sa);
        launch_rho(b, 0, RM_ADJ, b->q, b->mu, b->g, 0, 0, nullptr);
        launch_back(b, 0, 3, nullptr, nullptr, b->edel, nullptr, nullptr, nullptr);
        hipLaunchKernelGGL(kb_adj_rhs, dim3(bdiv(n, 256)), dim3(256), 0, b->st, n, b->type, b->tab_d, b->sdir,
                           b->efg_rec, b->g, b->edel, b->srhs, b->rmsd_i);
        double *s2 = T + 8, *r2 = T + 8 + nc;
        segsum(b, b->rmsd_i, n, 1, 0, s2);
        // (pol^-1 - T) lam = s from lam = pol s: conjugate gradients per
        // configuration (symmetric pol, as the SCF of the mixed form, in FP64),
        // then Jacobi sweeps until the recomputed residual meets the test
        // (one sweep after a converged CG: its residual is carried, not
        // recomputed)
        hipLaunchKernelGGL(kb_adj_start, dim3(bdiv(n, 256)), dim3(256), 0, b->st, n, b->type, b->tab_d, b->srhs,
                           b->lam);
        hipLaunchKernelGGL(kb_set_conf_int, dim3(bdiv(nc, 256)), dim3(256), 0, b->st, nc, b->act, 1);
        hipLaunchKernelGGL(kb_set_conf_int, dim3(bdiv(nc, 256)), dim3(256), 0, b->st, nc, b->iscf, 0);
        const int maxit = 1000;
        const double tol = 1e-13;
        int launched = 0;
        if (b->cg_sym) {
            const dim3 g1(bdiv(n, 256)), g3(bdiv(3 * n, 256));
            induced_field(b, 0, b->lam, nullptr, b->eT, b->act);
            hipLaunchKernelGGL(kb_add3, g3, dim3(256), 0, b->st, n3, b->srhs, b->eT, b->cg_e);
            hipLaunchKernelGGL(kb_cg_start, g1, dim3(256), 0, b->st, n, b->cfg, b->act, b->type, b->tab_d, b->cg_e,
                               b->lam, b->cg_r, b->cg_z, b->cg_p, b->cg_part, 1);
            segsum(b, b->cg_part, n, 2, 0, b->cg_s);
            hipLaunchKernelGGL(kb_cg_control, dim3(1), dim3(1024), 0, b->st, nc, b->conf, b->cg_s, 1, b->act,
                               b->iscf, b->rmsd, b->cg_rz, b->cg_beta, maxit, 0, tol, b->any, (const double *)s2);
            int done = 1;
            while (done < maxit) {
                if ((rc = read_any(b))) return rc;
                if (!*b->any_h) break;
                for (int k = 0; k < 4 && done < maxit; k++, done++) {
                    induced_field(b, 0, b->cg_p, nullptr, b->cg_w, b->act);
                    hipLaunchKernelGGL(kb_cg_pq, g1, dim3(256), 0, b->st, n, b->cfg, b->act, b->type, b->tab_d,
                                       b->cg_p, b->cg_w, b->cg_q, b->cg_part);
                    segsum(b, b->cg_part, n, 1, 0, b->cg_pq);
                    hipLaunchKernelGGL(kb_cg_update, g1, dim3(256), 0, b->st, n, b->cfg, b->act, b->type, b->tab_d,
                                       b->cg_rz, b->cg_pq, b->cg_p, b->cg_q, b->cg_w, b->lam, b->cg_r, b->cg_e,
                                       b->cg_z, b->cg_part, 1);
                    segsum(b, b->cg_part, n, 2, 0, b->cg_s);
                    hipLaunchKernelGGL(kb_cg_control, dim3(1), dim3(1024), 0, b->st, nc, b->conf, b->cg_s, 0, b->act,
                                       b->iscf, b->rmsd, b->cg_rz, b->cg_beta, maxit, 0, tol, b->any,
                                       (const double *)s2);
                    hipLaunchKernelGGL(kb_cg_dir, g3, dim3(256), 0, b->st, n, b->cfg, b->act, b->cg_beta, b->cg_z,
                                       b->cg_p);
                }
            }
            hipLaunchKernelGGL(kb_set_conf_int, dim3(bdiv(nc, 256)), dim3(256), 0, b->st, nc, b->act, 1);
        }
        while (launched < maxit) {
            for (int k = 0; k < 4 && launched < maxit; k++, launched++) {
                induced_field(b, 0, b->lam, nullptr, b->eT, b->act);
                hipLaunchKernelGGL(kb_adj_sweep, dim3(bdiv(n, 256)), dim3(256), 0, b->st, n, b->cfg, b->act, b->type,
                                   b->tab_d, b->srhs, b->eT, b->lam, b->rmsd_i);
                segsum(b, b->rmsd_i, n, 1, 0, r2);
                hipLaunchKernelGGL(kb_adj_control, dim3(1), dim3(1024), 0, b->st, nc, r2, s2, b->act, b->iscf, b->rmsd,
                                   maxit, tol, b->any);
            }
            if ((rc = read_any(b))) return rc;
            if (!*b->any_h) break;
        }
        std::vector<int> it(nc);
        std::vector<double> rs(nc);
        BCHK(hipMemcpy(it.data(), b->iscf, sizeof(int) * nc, hipMemcpyDeviceToHost));
        BCHK(hipMemcpy(rs.data(), b->rmsd, sizeof(double) * nc, hipMemcpyDeviceToHost));
        for (int c = 0; c < nc; c++) {
            iters_max = std::max(iters_max, it[c]);
            resid_max = std::max(resid_max, rs[c]);
        }
        // damping and polarisability terms
        if (b->ldmp) {
            a.out = cols + (size_t)4 * nt * n;
            adj_kernel(b, 1, a, sa);
        }
        hipLaunchKernelGGL(kb_adj_pol, dim3(bdiv(n, 256)), dim3(256), 0, b->st, n, b->type, b->tab_d, cu, b->lam,
                           b->mu, cols + (size_t)6 * nt * n);
    }
    double *TS = T + 8 + 2 * (size_t)nc;
    hipLaunchKernelGGL(kb_typesum, dim3(nt, NCOL), dim3(256), 0, b->st, cols, (size_t)n, NCOL, n, b->type, TS);
    std::vector<double> h((size_t)nt * NCOL);
    BCHK(hipMemcpyAsync(h.data(), TS, sizeof(double) * h.size(), hipMemcpyDeviceToHost, b->st));
    BCHK(hipStreamSynchronize(b->st));
    // per-type sums -> gradient (short range symmetrised over the pair)
    for (int ti = 0; ti < nt; ti++)
        for (int tj = 0; tj < nt; tj++) {
            double v[4];
            for (int k = 0; k < 4; k++) {
                double x = h[(size_t)ti * NCOL + 4 * tj + k];
                if (ti != tj) x += h[(size_t)tj * NCOL + 4 * ti + k];
                v[k] = x;
            }
            gr->dA[ti][tj] = v[0]; gr->dB[ti][tj] = v[1]; gr->dC[ti][tj] = v[2]; gr->dD[ti][tj] = v[3];
            gr->ddamp_b[ti][tj] = h[(size_t)ti * NCOL + 4 * nt + 2 * tj];
            gr->ddamp_c[ti][tj] = h[(size_t)ti * NCOL + 4 * nt + 2 * tj + 1];
        }
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 9; k++) gr->dpol[t][k / 3][k % 3] = h[(size_t)t * NCOL + 6 * nt + k];
    gr->adjoint_iterations = iters_max;
    gr->adjoint_residual = resid_max;
    (void)n3;
    return rc == MDFF_ERR_SCF ? MDFF_ERR_SCF : MDFF_OK;
}

// ------------------------------------------------------------ tuning
int mdff_batch_set_tuning(mdff_batch *b, int lanes, int kpb, int ipb, int bpc)
{
    if (!b) return MDFF_ERR_ARG;
    if (lanes) { if (lanes != 16 && lanes != 32 && lanes != 64) return MDFF_ERR_ARG; b->lanes = lanes; }
    if (kpb) { if (kpb != 32 && kpb != 64 && kpb != 128) return MDFF_ERR_ARG; b->kpb = kpb; }
    // ions per back-projection item: 16 or 32 (matrix-core sums), 64, 128 or
    // 256 (vector-ALU sums)
    if (ipb) {
        if (ipb == 16 || ipb == 32) b->mfr = ipb / 16;
        else if (ipb == 64 || ipb == 128 || ipb == 256) b->ipb = ipb;
        else return MDFF_ERR_ARG;
    }
    if (bpc) { if (bpc < 1 || bpc > 32) return MDFF_ERR_ARG; b->bpc = bpc; }
    b->tuned = 1;
    int rc = build_items(b);
    if (rc) return rc;
    BCHK(hipDeviceSynchronize());
    return MDFF_OK;
}

// time `iters` runs of fn between events on the batch stream
template <typename Fn>
static int time_it(mdff_batch *b, Fn fn, int iters, double *ms)
{
    fn();
    BCHK(hipEventRecord(b->ev[2], b->st));
    for (int k = 0; k < iters; k++) fn();
    BCHK(hipEventRecord(b->ev[3], b->st));
    BCHK(hipEventSynchronize(b->ev[3]));
    float t = 0.0f;
    BCHK(hipEventElapsedTime(&t, b->ev[2], b->ev[3]));
    *ms = (double)t / iters;
    return MDFF_OK;
}

int mdff_batch_autotune(mdff_batch *b, int out[4])
{
    if (!b) return MDFF_ERR_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return MDFF_ERR_DEVICE;
    int n = b->n, rc;
    int lo = lo_b(b);
    // state: list, tables, charges
    if ((rc = build_list(b))) return rc;
    if (b->p.lcoulombic) {
        hipLaunchKernelGGL(kb_eikr<double>, dim3(bdiv(n, 128)), dim3(128), 0, b->st, n, b->conf, b->cfg, b->r, b->TS,
                           b->by, b->bz, b->tab64, b->tabT64);
        if (lo) hipLaunchKernelGGL(kb_eikr<float>, dim3(bdiv(n, 128)), dim3(128), 0, b->st, n, b->conf, b->cfg, b->r,
                                   b->TS, b->by, b->bz, b->tab32, b->tabT32);
    }
    double best;
    // real-space lanes: the Ewald force kernel
    {
        int cand[3] = {16, 32, 64}, pick = b->lanes;
        best = 1e30;
        for (int L : cand) {
            b->lanes = L;
            double ms;
            DirArgs a = bdir_args(b, b->q, b->mu, b->ldmp, 0, nullptr);
            if ((rc = time_it(b, [&]() { launch_bdir(b, a, lo, 1); }, 5, &ms))) return rc;
            if (ms < best) { best = ms; pick = L; }
        }
        b->lanes = pick;
    }
    if (b->p.lcoulombic && b->nktot > 0 && b->mf) {
        // matrix-core sums: 16 or 32 ions per back-projection item (the
        // forward's items, 32 box points of one configuration, are fixed)
        int pr = b->mfr;
        best = 1e30;
        for (int r : {1, 2}) {
            b->mfr = r;
            if ((rc = build_items(b))) return rc;
            double ms;
            if ((rc = time_it(b, [&]() {
                     launch_back(b, lo, 9, b->q, b->mu, b->ef_rec, b->efg_rec, b->f_rec, nullptr);
                 }, 5, &ms)))
                return rc;
            if (ms < best) { best = ms; pr = r; }
        }
        b->mfr = pr;
        if ((rc = build_items(b))) return rc;
    } else if (b->p.lcoulombic && b->nktot > 0) {
        int pk = b->kpb, pi = b->ipb, pb = b->bpc;
        best = 1e30;
        for (int bpc : {2, 4, 8}) {
            for (int kpb : {32, 64, 128}) {
                b->kpb = kpb; b->bpc = bpc;
                if ((rc = build_items(b))) return rc;
                double ms;
                if ((rc = time_it(b, [&]() { launch_rho(b, lo, RM_QMU, b->q, b->mu, nullptr, 1, 1, nullptr); }, 5, &ms)))
                    return rc;
                if (ms < best) { best = ms; pk = kpb; pb = bpc; }
            }
        }
        b->kpb = pk; b->bpc = pb;
        best = 1e30;
        for (int ipb : {64, 128, 256}) {
            b->ipb = ipb;
            if ((rc = build_items(b))) return rc;
            double ms;
            if ((rc = time_it(b, [&]() {
                     launch_back(b, lo, 9, b->q, b->mu, b->ef_rec, b->efg_rec, b->f_rec, nullptr);
                 }, 5, &ms)))
                return rc;
            if (ms < best) { best = ms; pi = ipb; }
        }
        b->ipb = pi;
        if ((rc = build_items(b))) return rc;
    }
    b->tuned = 1;
    if (out) {
        out[0] = b->lanes;
        out[1] = b->mf ? 32 : b->kpb;
        out[2] = b->mf ? 16 * b->mfr : b->ipb;
        out[3] = b->bpc;
    }
    return MDFF_OK;
}

int mdff_batch_get_ewald(mdff_batch *b, int c, double *alpha, int kES[3], int *nk)
{
    if (!b || c < 0 || c >= b->nconf) return MDFF_ERR_ARG;
    const BConfD &K = b->conf_h[c];
    if (alpha) *alpha = K.alpha;
    if (kES) { kES[0] = K.Kx; kES[1] = K.Ky; kES[2] = K.Kz; }
    if (nk) *nk = K.nk;
    return MDFF_OK;
}

size_t mdff_batch_sizeof_grad(void) { return sizeof(mdff_grad); }

int mdff_batch_get_timing(mdff_batch *b, double *ms, long long *sweeps)
{
    if (!b) return MDFF_ERR_ARG;
    if (ms) *ms = b->last_ms;
    if (sweeps) *sweeps = b->sweeps;
    return MDFF_OK;
}
