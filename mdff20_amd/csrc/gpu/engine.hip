// engine.hip -- context, orchestration and C ABI of the MI355X engine for
// mdff20's force and integration path.  One HIP stream per context; state is
// device resident; the host synchronises only to read the SCF convergence
// flag, the neighbour-list overflow flag and observables.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "kernels.h"
#include "md_kernels.h"
#include "mfma_kernels.h"
#include "gemm_kernels.h"
#include "typetab.h"
#include "mdff_gpu.h"
#include "mdff_host.h"
#include "mdff_rdf.h"
#include "engine_view.h"

#define HIPCHK(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            fprintf(stderr, "mdff: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            return MDFF_ERR_DEVICE;                                                               \
        }                                                                                         \
    } while (0)

namespace {

template <typename T>
int dalloc(T **p, size_t n)
{
    if (n == 0) n = 1;
    if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) return MDFF_ERR_NOMEM;
    if (hipMemset(*p, 0, n * sizeof(T)) != hipSuccess) return MDFF_ERR_DEVICE;
    return MDFF_OK;
}

template <typename T>
void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

struct mdff_ctx {
    mdff_params p;
    mdff_units u;
    mdff_cell cell;
    CellD cd;
    int n = 0, ntype = 0, precision = MDFF_FP64, device = 0;
    hipStream_t st = nullptr;
    hipStream_t st2 = nullptr;       // real-space kernels overlapping the k-space sums (default priority)
    hipEvent_t xev[2] = {nullptr, nullptr};   // fork (st -> st2) and join (st2 -> st)
    // derived (field.c:186-335, md.c:120-141)
    int lqch = 0, ldip_static = 0, lpol = 0, ldmp = 0, lpim = 0, lnonbonded = 0;
    int ldip_sticky = 0;  // coulombic.c:115
    int integ = MDFF_NVE_VV, rescale_allowed = 0;
    double srcutsq = 0, lrcutsq = 0, alpha = 0;
    int kES[3] = {0, 0, 0};
    double kT = 0, dt = 0, tauT = 0, timesca = 0;
    double selfa = 0, selfa2 = 0, ttpiV = 0, tfpiV = 0;
    int npol = 0;
    std::vector<int> typ;
    std::vector<double> mass;
    TypeTab tab;
    // device state (SoA, 3n unless stated)
    double *r = nullptr, *v = nullptr, *f = nullptr, *fold = nullptr, *rs = nullptr, *s = nullptr, *s0 = nullptr;
    int *type = nullptr;
    double *invmass = nullptr, *massd = nullptr, *q = nullptr, *qz = nullptr;
    double *mu_static = nullptr, *mu_ind = nullptr, *mu_tot = nullptr, *hist = nullptr;
    double *ef_stat = nullptr, *ef = nullptr, *ef_dir = nullptr, *ef_rec = nullptr, *efg = nullptr,
           *efg_dir = nullptr, *efg_rec = nullptr;
    double *f_sr = nullptr, *f_dir = nullptr, *f_rec = nullptr;
    double *acc_sr = nullptr;  // 14n
    double *uw_dir = nullptr;  // 7n
    double *selfbuf = nullptr; // 2n
    double *ke = nullptr;      // n
    double *upol_i = nullptr, *rmsd_i = nullptr;  // n
    double *tmp_q = nullptr, *tmp_mu = nullptr, *tmp_qraw = nullptr, *tmp_mur = nullptr;  // drop-in scratch
    double *dc_R = nullptr, *dc_d = nullptr, *dc_g = nullptr;  // anchored CG: residual, step, T p
    double *cg_z = nullptr, *cg_p = nullptr, *cg_part = nullptr; // preconditioned residual, direction, sums
    double *ko_f = nullptr, *ko_dmu = nullptr; // scfkO: residual field and inner-loop dipoles
    double *t2ko = nullptr;   // scfkO: dipole tensors of the polarisable pairs at alpha 0.001
    double *scal = nullptr;
    int *ctl = nullptr;
    TypeTab *tab_d = nullptr;
    int *pidx = nullptr;
    // stored dipole tensors of the SCF real space (k_t2_build / k_t2_field)
    int *pinv = nullptr, *cntp = nullptr, *idxp = nullptr, *t2_ovf = nullptr;
    double *t2 = nullptr;
    float *t2f = nullptr;
    int capP = 0, t2_ok = 0;
    // neighbour list
    unsigned *nbr = nullptr;
    int *cnt = nullptr;
    int cap = 0, have_list = 0;
    // k space
    int Kx = 0, Ky = 0, Kz = 0, P = 0;
    size_t nkf = 0;
    int nk = 0;
    double *kvec = nullptr, *Ak = nullptr, *kcoe = nullptr, *W = nullptr, *ek = nullptr;
    // e^{i n theta} tables [lo][polar][axis]: lo = FP32 copy (mixed form),
    // polar = polarisable-ion subset (SCF)
    void *tab3[2][2][3] = {{{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}},
                           {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}}};
    double *kpart = nullptr, *bpart = nullptr;
    double *ksum = nullptr;   // [nkf][8] chunk-summed forward partials of the matrix-core path
    int fwd_part = 0;         // last forward launch: 0 partials in kpart (k_sk_reduce sums them), 1 summed into ksum,
                              // 2 FP32 partials of a matrix-core sum left in kpart for k_psum_reduce_cm
    bool psum_defer = false;  // set by rec_part: the split-FP16 forward leaves its partials (fwd_part 2)
    int psum_ch = 3;          // with psum_defer: moments (3 or 6) of the k_psum_reduce_cm that follows
    int no_fuse = 0;          // tuning: separate partial sums, reduction and pair-tensor finish
    int fwd_nch = 0;          // fwd_part 2: K ranges of the partials
    const double *fwd_inv = nullptr;   // fwd_part 2: channel scales to undo
    bool fwd_comb = false;             // fwd_part 2: partials hold rho = i k.S_mu (one channel)
    int nchunk = 0, nsplit = 0, chunk = 0, rows_per_split = 0;
    int nchunk_p = 0, chunk_p = 0;
    // tuning
    int lanes = 16;           // lanes per ion of the pair kernels and the pair-tensor field
    int use_mfma = 1;         // S(k) forward and back-projection on the matrix cores
    int g_cw[2] = {1, 2};     // column tiles per wave of the forward GEMM (FP64, FP32/split-FP16)
    int g_fblocks = 0;        // target blocks of the forward GEMM (0: one round of resident blocks)
    int g_bblocks = 0;        // target blocks of the back-projection GEMM (0: one round)
    int ncu = 256;            // compute units of the device
    int nselp[2] = {0, 0};    // ions padded to 128 (all, polarisable)
    int Pp = 0, Pk = 0;       // (ny,nz) points padded: forward rows, back-projection K
    void *yz[4][2][2] = {};   // [mode][polar][hi,lo]: YZ[p][2j+ri]
    void *yzt[4][2][2] = {};  // [mode][polar][hi,lo]: YZt[j][2p+ri]
    int yz_ok[4][2] = {};     // stored operands match the current tables
    void *aw[4][2] = {};      // [mode][hi,lo]: rows (nx,c,re/im) of the forward sum
    void *wm[4][2] = {};      // [mode][hi,lo]: columns (nx,m,re/im) of the back sum
    double *gsc = nullptr;    // split-FP16 scales: weight inverses [8,16), WM column inverses from 16
    double *gwp = nullptr;    // split-FP16 per-block weight channel maxima [4][n/256]
    double *gwq = nullptr;    // the same for the charges alone (all ions), formed once [n/256]
    int gwq_ok = 0;
    unsigned long long *gcm = nullptr;   // split-FP16 WM column maxima [(Kx+1) 6] (double bits)
    double *chk_part = nullptr;          // SCF rmsd / u_pol partial sums [2][n/256]
    double *pq_buf = nullptr;            // CG p.q partial sums, one per fin_t2 block [n/4]
    double *t2dir = nullptr;             // pair-tensor field of a CG direction (back-projection's extra rows) [3n]
    struct GTune { int mode, dir, tiles, polar, blocks; };
    std::vector<GTune> gtune;  // block targets of the stored-operand sums, timed on first use
    int mf_lo = KM_F16X3;     // matrix-core mode of the FP32-table evaluations (static field, forces)
    int mf_inner = KM_F16X3;  // mode of the mixed-form low-precision SCF steps
    int mf_anchor = KM_F16X3; // anchored CG: anchor evaluation in FP64 (0) or this mode over the FP32 tables
    int wpart_ready = 0;      // gwp holds the channel maxima of the next split-FP16 forward sum's weights
    int wpart_nb = 0;         // with wpart_ready: partials per channel in gwp (0: one per 256 selected ions)
    int tot_done = 0;         // k_mu_tot formed mu_tot, the self terms and (2: also) gwp for the force call
    int kin_fresh = 0;        // ke and S_KIN hold the current velocities' kinetic energy (within one run call)
    CellGrid cg{};                 // cell grid of the neighbour build
    int cells_set = 0;             // 1 grid chosen (cells_on tells whether it is used)
    int cells_on = 0;              // the list is built from the cell grid (else all pairs)
    int list_method = 0;           // 0 cells where the box allows, 1 all pairs always
    double *vc_part = nullptr;     // k_vcheck's per-block two largest displacements [2][n/256]
    unsigned *vc_cnt = nullptr;    // k_vcheck's finished-block count
    unsigned *chk_cnt = nullptr;   // k_scf_check's finished-block count
    int mf_blocks = 512;      // target blocks of the MFMA kernels
    int nsplit_max = 1;       // (ny,nz) splits the bpart buffer holds
    // host pinned mirrors
    int *ctl_h = nullptr;
    double *scal_h = nullptr;
    // NHC constants
    double nhcQ[16];
    double yosh[7];
    // timing
    hipEvent_t ev[8];
    hipEvent_t *probe_ev = nullptr; // mdff_time_kernel: events around a k-space GEMM launch
    int probe_k = 0, probe_max = 0;
    int probe_rep = 1;        // the GEMM launched this many times back to back between the events
    int have_events = 0;
    double t_ms[6] = {0, 0, 0, 0, 0, 0};
    int last_nscf = 0;
    int hint_fp64 = 0, hint_cg = 0, hint_ko = 0; // iterations of the last SCF of each solver
    int hint_ref = 0;         // correction steps of the last FP64 refinement
    double last_rmsd = 0;
    int scf_check = 0;        // mixed form: FP64 residual of the SCF every scf_check-th force call (0 never)
    int scf_refine = 0;       // mixed form, anchored CG: FP64 refinement of the converged dipoles
    // device trajectory: frames of r, v, f (9 n doubles each, SoA) at the
    // TRAJFF cadence of run_md (md.c:275-277, iopsnode io.c:67-69)
    int traj_every = 0, traj_start = 0, traj_npas = 0, traj_cap = 0, traj_n = 0, traj_drop = 0;
    double *traj = nullptr;
    std::vector<int> traj_step;
    int chk_done = 0;         // the last force call evaluated it (S_RMSD_CHK holds the sum)
    int scf_chunk = 4;
    int cg_first = 8;         // anchored CG: steps launched before a context's first convergence poll
    int scf_solver = 2;       // mixed form: 0 the reference's Jacobi schedule (FP64 fields), 2 anchored CG
    int fp64_cg = 0;          // FP64 form: anchored CG with FP64 steps instead of the reference's Jacobi schedule
    int cg_ok = 0;            // every polarisability symmetric positive definite (CG applies)
    long long n_force = 0, n_scf_fp64 = 0, n_scf_inner = 0;
};

// ------------------------------------------------------------------- helpers
static inline int lo_of(const mdff_ctx *c) { return c->precision == MDFF_MIXED ? 1 : 0; }

static int sync_ctl(mdff_ctx *c)
{
    HIPCHK(hipMemcpyAsync(c->ctl_h, c->ctl, sizeof(int) * C_COUNT, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return MDFF_OK;
}

static int sync_scal(mdff_ctx *c)
{
    // scal and ctl are one allocation on each side: one copy
    HIPCHK(hipMemcpyAsync(c->scal_h, c->scal, sizeof(double) * S_COUNT + sizeof(int) * C_COUNT,
                          hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return MDFF_OK;
}

// mdff_time_kernel brackets each k-space GEMM launch with a pair of events
// launches of a k-space GEMM between its probe events
static int probe_reps(const mdff_ctx *c, bool probe)
{
    return (probe && c->probe_ev && c->probe_k < c->probe_max) ? c->probe_rep : 1;
}

static void probe_mark(mdff_ctx *c, int end)
{
    if (!c->probe_ev || c->probe_k >= c->probe_max) return;
    (void)hipEventRecord(c->probe_ev[2 * c->probe_k + end], c->st);
    if (end) c->probe_k++;
}

static void set_int(mdff_ctx *c, int slot, int val)
{
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->st, c->ctl + slot, val);
}

// lanes per polarisable ion of k_sk_back_fin_t2 (8, 16, 32 or 64; at most 64,
// which sizes pq_buf)
static int fin_t2_lanes(const mdff_ctx *c) { return c->lanes; }

static void colsum(mdff_ctx *c, const double *in, int n, int ld, int ncols, double *out, const int *active = nullptr)
{
    hipLaunchKernelGGL(k_colsum, dim3(ncols), dim3(1024), 0, c->st, in, n, ld, out, active);
}

// Second stream: the real-space pair kernels (vector ALU) run beside the
// k-space sums (matrix cores).  aux_fork makes st2 wait for everything
// issued on st so far; aux_join makes st wait for everything on st2.
// OnAux sends the launches of its scope (all helpers launch on c->st) to st2.
static void aux_fork(mdff_ctx *c)
{
    if (!c->st2) return;
    (void)hipEventRecord(c->xev[0], c->st);
    (void)hipStreamWaitEvent(c->st2, c->xev[0], 0);
}

static void aux_join(mdff_ctx *c)
{
    if (!c->st2) return;
    (void)hipEventRecord(c->xev[1], c->st2);
    (void)hipStreamWaitEvent(c->st, c->xev[1], 0);
}

struct OnAux {
    mdff_ctx *c;
    hipStream_t keep;
    explicit OnAux(mdff_ctx *cc) : c(cc), keep(cc->st) { if (cc->st2) cc->st = cc->st2; }
    ~OnAux() { c->st = keep; }
};

// ------------------------------------------------------------- type tables
static void build_tables(mdff_ctx *c)
{
    typetab_build(c->p, c->ntype, c->srcutsq, c->tab);
    // the anchored CG needs pol^-T - T symmetric positive definite: each
    // polarisability symmetric with positive leading minors
    c->cg_ok = 1;
    for (int t = 0; t < c->ntype; t++) {
        if (!c->tab.polar[t]) continue;
        const double *m = c->tab.pol[t];
        double sc = fabs(m[0]) + fabs(m[4]) + fabs(m[8]);
        bool sym = fabs(m[1] - m[3]) <= 1e-14 * sc && fabs(m[2] - m[6]) <= 1e-14 * sc && fabs(m[5] - m[7]) <= 1e-14 * sc;
        double m2 = m[0] * m[4] - m[1] * m[3];
        double m3 = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                    m[2] * (m[3] * m[7] - m[4] * m[6]);
        if (!(sym && m[0] > 0.0 && m2 > 0.0 && m3 > 0.0)) c->cg_ok = 0;
    }
}

// ------------------------------------------------------------ neighbour list
// The cell grid of the neighbour build, chosen once per context: per axis
// nc cells of at least rmax / 2 across (rmax: the larger list cutoff, skin
// included), so pairs within rmax are at most reach = 2 cells apart; with
// fewer than 2 reach + 1 cells on an axis, cells of at least rmax (reach 1);
// with fewer than 3 of those the box is too small for a grid and every pair
// is scanned.  Widths are the perpendicular ones, 1/|B_d|, so triclinic
// cells are covered.
static bool grid_dims(const mdff_ctx *c, CellGrid &g)
{
    if (c->list_method == 1 || !c->p.lverletL) return false;
    const double rmax = std::max(c->p.cutshortrange, c->p.cutlongrange) + c->p.skindiff;
    double w[3];
    for (int d = 0; d < 3; d++) {
        const double *b = c->cell.B[d];
        w[d] = 1.0 / sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    }
    int nc[3], reach[3];
    bool ok = false;
    for (int sub = 2; sub >= 1 && !ok; sub--) {
        ok = true;
        for (int d = 0; d < 3; d++) {
            nc[d] = (int)floor(sub * w[d] / rmax);
            reach[d] = (int)ceil(nc[d] * rmax / w[d] - 1e-9);
            if (nc[d] < 1 || reach[d] < 1 || nc[d] < 2 * reach[d] + 1) ok = false;
        }
    }
    if (!ok) return false;
    const long long ncell = (long long)nc[0] * nc[1] * nc[2];
    const int nst = (2 * reach[0] + 1) * (2 * reach[1] + 1) * (2 * reach[2] + 1);
    if (ncell > (1 << 22) || nst > 125) return false;
    // a stencil that covers most of the grid scans about every pair anyway:
    // the all-pairs build is cheaper there unless the grid is asked for
    if (c->list_method == 0 && ncell < 4LL * nst) return false;
    for (int d = 0; d < 3; d++) {
        g.nc[d] = nc[d];
        g.reach[d] = reach[d];
    }
    g.ncell = (int)ncell;
    g.nst = nst;
    return true;
}

static int setup_cells(mdff_ctx *c)
{
    c->cells_set = 1;
    c->cells_on = 0;
    CellGrid &g = c->cg;
    if (!grid_dims(c, g)) return MDFF_OK;
    int rc;
    if ((rc = dalloc(&g.ccount, (size_t)g.ncell)) || (rc = dalloc(&g.cstart, (size_t)g.ncell + 1)) ||
        (rc = dalloc(&g.cidx, (size_t)c->n)) || (rc = dalloc(&g.csort, (size_t)c->n)) ||
        (rc = dalloc(&g.cell, (size_t)c->n)) || (rc = dalloc(&g.slot, (size_t)c->n)))
        return rc;
    c->cells_on = 1;
    return MDFF_OK;
}

static void free_cells(mdff_ctx *c)
{
    dfree(c->cg.ccount);
    dfree(c->cg.cstart);
    dfree(c->cg.cidx);
    dfree(c->cg.csort);
    dfree(c->cg.cell);
    dfree(c->cg.slot);
    c->cells_set = c->cells_on = 0;
}

// the list build after k_vcheck (gated on its rebuild flag): cell grid, or
// every pair; g is the grid (its arrays may be a caller's scratch)
static void launch_list_build(mdff_ctx *c, const CellGrid &g, bool cells, int all, double csr, double clr)
{
    int n = c->n;
    if (cells && !all) {
        hipLaunchKernelGGL(k_cell_bin, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->s, g, c->ctl);
        hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, c->st, g, c->ctl);
        hipLaunchKernelGGL(k_cell_fill, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, g, c->ctl);
        hipLaunchKernelGGL(k_cell_rank, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, g, c->ctl);
        hipLaunchKernelGGL(k_build_cells, dim3(g.ncell), dim3(256), 0, c->st, n, c->cd, c->s, g, csr * csr,
                           clr * clr, c->cap, c->nbr, c->cnt, c->s0, c->ctl);
        hipLaunchKernelGGL(k_row_sort, dim3(cdiv(n, 4)), dim3(256), 0, c->st, n, c->cap, c->nbr, c->cnt, c->ctl);
    } else {
        hipLaunchKernelGGL(k_build_allpairs, dim3(cdiv(n, 4)), dim3(256), 0, c->st, n, c->cd, c->s, csr * csr,
                           clr * clr, all, c->cap, c->nbr, c->cnt, c->s0, c->ctl);
    }
}

static int build_list(mdff_ctx *c, int force)
{
    int n = c->n;
    if (!c->cells_set) {
        int rc = setup_cells(c);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_frac, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->cd, c->r, c->r + n, c->r + 2 * n, c->s);
    double csr = c->p.cutshortrange + c->p.skindiff, clr = c->p.cutlongrange + c->p.skindiff;
    int all = !c->p.lverletL;
    for (int attempt = 0; attempt < 3; attempt++) {
        if (all) {
            // no Verlet list: every pair is a candidate, independent of positions
            set_int(c, C_OVERFLOW, 0);
            set_int(c, C_MAXCNT, 0);
            set_int(c, C_REBUILD, (force || attempt > 0 || !c->have_list) ? 1 : 0);
        } else {
            if (force || attempt > 0) set_int(c, C_FIRST, 1);
            // the check also clears the overflow flag and count of the build
            hipLaunchKernelGGL(k_vcheck, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->cd, c->s, c->s0,
                               c->p.skindiff, c->ctl, 1, c->vc_part, c->vc_cnt);
            if (attempt > 0) hipLaunchKernelGGL(k_add_int, dim3(1), dim3(1), 0, c->st, c->ctl + C_NUPD, -1);
        }
        // the build only when the check asks for it (the check's flag is
        // read back first; most steps end here)
        int rc = sync_ctl(c);
        if (rc) return rc;
        if (!c->ctl_h[C_REBUILD]) {
            c->have_list = 1;
            return MDFF_OK;
        }
        launch_list_build(c, c->cg, c->cells_on, all, csr, clr);
        rc = sync_ctl(c);
        if (rc) return rc;
        if (c->ctl_h[C_OVERFLOW] == 0) {
            c->have_list = 1;
            return MDFF_OK;
        }
        // grow the capacity and build again at the same positions
        int need = c->ctl_h[C_MAXCNT];
        int ncap = ((need * 5) / 4 + 64 + 63) / 64 * 64;
        if (ncap > ((n + 63) / 64) * 64) ncap = ((n + 63) / 64) * 64;
        dfree(c->nbr);
        // no zero fill: the build writes every entry below each ion's count
        if (hipMalloc((void **)&c->nbr, sizeof(unsigned) * (size_t)n * ncap) != hipSuccess) {
            c->nbr = nullptr;
            return MDFF_ERR_NOMEM;
        }
        c->cap = ncap;
    }
    return MDFF_ERR_LIST_OVERFLOW;
}

// ----------------------------------------------------------------- dispatch
template <typename real, bool EF, bool EQ, bool EM, bool F>
static void launch_dir_t(mdff_ctx *c, const DirArgs &a)
{
    int L = c->lanes;
    unsigned blocks = cdiv((size_t)c->n * L, 256);
    if (L == 8) hipLaunchKernelGGL((k_ewald_dir<real, 8, EF, EQ, EM, F>), dim3(blocks), dim3(256), 0, c->st, a);
    else if (L == 16) hipLaunchKernelGGL((k_ewald_dir<real, 16, EF, EQ, EM, F>), dim3(blocks), dim3(256), 0, c->st, a);
    else if (L == 64) hipLaunchKernelGGL((k_ewald_dir<real, 64, EF, EQ, EM, F>), dim3(blocks), dim3(256), 0, c->st, a);
    else hipLaunchKernelGGL((k_ewald_dir<real, 32, EF, EQ, EM, F>), dim3(blocks), dim3(256), 0, c->st, a);
}

template <typename real>
static void launch_dir_r(mdff_ctx *c, const DirArgs &a, int ef, int eq, int em, int f)
{
    int code = (ef ? 8 : 0) | (eq ? 4 : 0) | (em ? 2 : 0) | (f ? 1 : 0);
    switch (code) {
    case 8: launch_dir_t<real, true, false, false, false>(c, a); break;
    case 1: launch_dir_t<real, false, false, false, true>(c, a); break;
    case 3: launch_dir_t<real, false, false, true, true>(c, a); break;
    case 9: launch_dir_t<real, true, false, false, true>(c, a); break;
    case 11: launch_dir_t<real, true, false, true, true>(c, a); break;
    case 14: launch_dir_t<real, true, true, true, false>(c, a); break;
    case 15: launch_dir_t<real, true, true, true, true>(c, a); break;
    case 6: launch_dir_t<real, false, true, true, false>(c, a); break;
    case 7: launch_dir_t<real, false, true, true, true>(c, a); break;
    case 10: launch_dir_t<real, true, false, true, false>(c, a); break;
    case 2: launch_dir_t<real, false, false, true, false>(c, a); break;
    case 0: break;
    default: launch_dir_t<real, true, true, true, true>(c, a); break;
    }
}

static void launch_dir(mdff_ctx *c, const DirArgs &a, int lo, int ef, int eq, int em, int f)
{
    if (lo) launch_dir_r<float>(c, a, ef, eq, em, f);
    else launch_dir_r<double>(c, a, ef, eq, em, f);
}

static DirArgs dir_args(mdff_ctx *c, const double *q, const double *mu, int lcouldamp, int q_zero,
                        const int *active)
{
    DirArgs a;
    memset(&a, 0, sizeof(a));
    a.n = c->n;
    a.cap = c->cap;
    a.c = c->cd;
    a.s = c->s;
    a.type = c->type;
    a.nbr = c->nbr;
    a.cnt = c->cnt;
    a.T = c->tab_d;
    a.q = q;
    a.mu = mu;
    a.alpha = c->alpha;
    a.rc2 = c->lrcutsq;
    a.lcouldamp = lcouldamp;
    a.active = active;
    a.q_zero = q_zero;
    a.ef = c->ef_dir;
    a.efg = c->efg_dir;
    a.f = c->f_dir;
    a.uw = c->uw_dir;
    return a;
}

template <typename real>
static void launch_eikr_t(mdff_ctx *c, const int *idx, int nsel, void *tx, void *ty, void *tz, const int *active,
                          void **tp)
{
    hipLaunchKernelGGL(k_eikr<real>, dim3(cdiv(nsel, 128)), dim3(128), 0, c->st, c->n, c->cd, c->Kx, c->Ky, c->Kz,
                       c->r, c->r + c->n, c->r + 2 * c->n, idx, nsel, (real *)tx, (real *)ty, (real *)tz, active,
                       tp ? c->pinv : nullptr, c->npol, tp ? (real *)tp[0] : nullptr, tp ? (real *)tp[1] : nullptr,
                       tp ? (real *)tp[2] : nullptr);
}

// the e^{in theta} tables of one ion set; for the all-ion set with
// with_polar also the polarisable-ion tables of the same precision (one pass)
static void launch_eikr(mdff_ctx *c, int polar, int lo, const int *active, bool with_polar = false)
{
    with_polar = with_polar && !polar && c->npol > 0;
    // stored operands built from these tables are stale now
    for (int ps = polar; ps <= (with_polar ? 1 : polar); ps++) {
        if (lo) { c->yz_ok[KM_F32][ps] = 0; c->yz_ok[KM_F16X3][ps] = 0; }
        else c->yz_ok[KM_F64][ps] = 0;
    }
    void **t = c->tab3[lo][polar];
    void **tp = with_polar ? c->tab3[lo][1] : nullptr;
    const int *idx = polar ? c->pidx : nullptr;
    int nsel = polar ? c->npol : c->n;
    if (lo) launch_eikr_t<float>(c, idx, nsel, t[0], t[1], t[2], active, tp);
    else launch_eikr_t<double>(c, idx, nsel, t[0], t[1], t[2], active, tp);
}

template <typename real, int C>
static void launch_fwd_t(mdff_ctx *c, const SkArgs &a, int nchunk)
{
    const int NY = 2 * c->Ky + 1, NZ = 2 * c->Kz + 1;
    size_t sh = (size_t)(2 * 32 + 2 * NY * 32 + 2 * NZ * 32 + C * 32) * sizeof(real);
    int ntile = (c->P + 255) / 256;
    hipLaunchKernelGGL((k_sk_fwd<real, C>), dim3((c->Kx + 1) * ntile, nchunk), dim3(256), sh, c->st, a);
}

template <typename real, int NCH>
static void launch_back_t(mdff_ctx *c, const BackArgs &a, int nsel)
{
    const int NZ = 2 * c->Kz + 1;
    size_t sh = (size_t)(2 * NZ * 64) * sizeof(real);
    hipLaunchKernelGGL((k_sk_back<real, NCH>), dim3(cdiv(nsel, 64), c->nsplit), dim3(64), sh, c->st, a);
}

static int launch_fwd_gemm(mdff_ctx *c, int polar, int lo, int C, const double *q_w, const double *mu_w,
                           const int *active, int *nchunk_out);
// a request to the back-projection for the pair-tensor field of mu at the
// polarisable ions (FP32 tensors) in extra grid rows of its launch; done is
// set when the launch took them
struct T2Rows {
    const float *t2;
    const double *mu;
    double *out;
    int done;
};
static int launch_back_gemm(mdff_ctx *c, int polar, int lo, int NCH, const int *active, int *nsplit_out,
                            bool wm_built, T2Rows *t2r);
static int wm_alloc(mdff_ctx *c, int mode);

// forward S(k) on the matrix cores; 0 when the shape is outside the GEMM
// kernels' tiles (or an operand buffer could not be allocated): the caller
// then takes the vector-ALU sum
static int launch_fwd_mfma(mdff_ctx *c, int polar, int lo, int C, const double *q_w, const double *mu_w,
                           const int *active, int *nchunk_out)
{
    int M = (c->Kx + 1) * C * 2;
    if (!c->use_mfma || M > MF_MT * 16) return 0;
    return launch_fwd_gemm(c, polar, lo, C, q_w, mu_w, active, nchunk_out);
}

// whether the back-projection GEMM takes a k box with CH moments: its
// column tiles (MB_CT) and the nx planes of its staged ex table (GB_KX1_MAX)
static bool g_back_fits(const mdff_ctx *c, int CH)
{
    return (c->Kx + 1) * CH * 2 <= MB_CT * 16 && c->Kx + 1 <= GB_KX1_MAX;
}

// back-projection on the matrix cores; returns 0 when outside the kernel's shapes
static int launch_back_mfma(mdff_ctx *c, int polar, int lo, int NCH, const int *active, int *nsplit_out,
                            bool wm_built = false, T2Rows *t2r = nullptr)
{
    if (!c->use_mfma || !g_back_fits(c, NCH == 9 ? 6 : 3)) return 0;
    return launch_back_gemm(c, polar, lo, NCH, active, nsplit_out, wm_built, t2r);
}

// ------------------------------------------------ stored-operand matrix-core sums
static size_t g_elt(int mode) { return mode == KM_F64 ? 8 : (mode == KM_F32 ? 4 : 2); }
static bool g_split_mode(int mode) { return mode == KM_F16X3; }

// allocation made while the context works: zeroed in stream order, since a
// null-stream memset is not ordered with the non-blocking context stream
static int galloc(mdff_ctx *c, void **p, size_t bytes)
{
    if (*p) return MDFF_OK;
    if (hipMalloc(p, bytes) != hipSuccess) { *p = nullptr; return MDFF_ERR_NOMEM; }
    if (hipMemsetAsync(*p, 0, bytes, c->st) != hipSuccess) return MDFF_ERR_DEVICE;
    return MDFF_OK;
}

// YZ and YZt of one table set in one mode, built once per table rebuild
static int ensure_yz(mdff_ctx *c, int mode, int polar, int lo)
{
    if (c->yz_ok[mode][polar]) return MDFF_OK;
    const int nsel = polar ? c->npol : c->n;
    const int nselp = c->nselp[polar];
    const size_t e = g_elt(mode);
    const size_t nb1 = (size_t)c->Pp * 2 * nselp * e, nb2 = (size_t)nselp * 2 * c->Pk * e;
    // YZt only for the FP64 mode's back-projection (the others form it)
    const bool wt = mode == KM_F64;
    int rc;
    if ((rc = galloc(c, &c->yz[mode][polar][0], nb1)) || (wt && (rc = galloc(c, &c->yzt[mode][polar][0], nb2))))
        return rc;
    if (g_split_mode(mode) && (rc = galloc(c, &c->yz[mode][polar][1], nb1))) return rc;
    const int NZ = 2 * c->Kz + 1;
    void **t = c->tab3[lo][polar];
    void *y0 = c->yz[mode][polar][0], *y1 = c->yz[mode][polar][1];
    void *z0 = c->yzt[mode][polar][0], *z1 = c->yzt[mode][polar][1];
    dim3 g1(nselp / 64, (std::max(c->Pp, c->Pk) + 31) / 32);
#define YZL(M, TT)                                                                                            \
    hipLaunchKernelGGL((k_yz_build<M, TT>), g1, dim3(256), 0, c->st, nsel, nselp, c->P, c->Pp, c->Pk, NZ,   \
                       (const TT *)t[1], (const TT *)t[2], y0, y1, z0, z1)
    dim3 g2(nselp / 64, (std::max(c->Pp, c->Pk) + 63) / 64);
#define YZ16(M)                                                                                               \
    hipLaunchKernelGGL(k_yz_build16<M>, g2, dim3(256), 0, c->st, nsel, nselp, c->P, c->Pp, c->Pk, NZ,          \
                       (const float *)t[1], (const float *)t[2], y0, y1, z0, z1)
    if (mode == KM_F64) YZL(KM_F64, double);
    else if (mode == KM_F32) YZ16(KM_F32);
    else YZ16(KM_F16X3);
#undef YZL
#undef YZ16
    c->yz_ok[mode][polar] = 1;
    return MDFF_OK;
}

template <int MODE, int CW>
static const void *gfwd_kernel_t(int MT)
{
    switch (g_mt_inst(MT)) {
    case 3: return (const void *)k_gemm_fwd<MODE, CW, 3>;
    case 5: return (const void *)k_gemm_fwd<MODE, CW, 5>;
    case 7: return (const void *)k_gemm_fwd<MODE, CW, 7>;
    default: return (const void *)k_gemm_fwd<MODE, CW, MF_MT>;
    }
}

template <int MODE>
static const void *gfwd_kernel_m(int CW, int MT)
{
    if (CW == 1) return gfwd_kernel_t<MODE, 1>(MT);
    if (CW == 4) return gfwd_kernel_t<MODE, 4>(MT);
    return gfwd_kernel_t<MODE, 2>(MT);
}

static const void *gfwd_kernel(int mode, int CW, int MT)
{
    if (mode == KM_F64) return gfwd_kernel_m<KM_F64>(CW, MT);
    if (mode == KM_F32) return gfwd_kernel_m<KM_F32>(CW, MT);
    return gfwd_kernel_m<KM_F16X3>(CW, MT);
}

template <int MODE, typename TT>
static const void *gback_kernel_t(int CT)
{
    switch (g_ct_inst(CT)) {
    case 4: return (const void *)k_gemm_back<MODE, 4, TT>;
    case 7: return (const void *)k_gemm_back<MODE, 7, TT>;
    case 13: return (const void *)k_gemm_back<MODE, 13, TT>;
    default: return (const void *)k_gemm_back<MODE, MB_CT, TT>;
    }
}

static const void *gback_kernel(int mode, int CT)
{
    if (mode == KM_F64) return gback_kernel_t<KM_F64, double>(CT);
    if (mode == KM_F32) return gback_kernel_t<KM_F32, float>(CT);
    return gback_kernel_t<KM_F16X3, float>(CT);
}

// launch the instantiation for the shape's tile count (GEN: B formed from
// the tables)
template <int MODE, int CW, bool GEN>
static void gfwd_mt(mdff_ctx *c, dim3 grid, int MT, const GFwdArgs &a)
{
    const int MTI = g_mt_inst(MT);
    const size_t lds = std::max(g_fwd_lds_bytes<MODE>(MTI, GEN), a.comb ? g_fwd_comb_lds_bytes(MTI, CW) : (size_t)0);
    if (MTI == 3) hipLaunchKernelGGL((k_gemm_fwd<MODE, CW, 3, GEN>), grid, dim3(256), lds, c->st, a);
    else if (MTI == 5) hipLaunchKernelGGL((k_gemm_fwd<MODE, CW, 5, GEN>), grid, dim3(256), lds, c->st, a);
    else if (MTI == 7) hipLaunchKernelGGL((k_gemm_fwd<MODE, CW, 7, GEN>), grid, dim3(256), lds, c->st, a);
    else hipLaunchKernelGGL((k_gemm_fwd<MODE, CW, MF_MT, GEN>), grid, dim3(256), lds, c->st, a);
}

template <int MODE, bool GEN = false>
static void gfwd_cw(mdff_ctx *c, dim3 grid, int CW, int MT, const GFwdArgs &a)
{
    if (CW == 1) gfwd_mt<MODE, 1, GEN>(c, grid, MT, a);
    else if (CW == 4) gfwd_mt<MODE, 4, GEN>(c, grid, MT, a);
    else gfwd_mt<MODE, 2, GEN>(c, grid, MT, a);
}

template <int MODE, typename TT>
static void gback_ct(mdff_ctx *c, dim3 grid, int CT, const GBackArgs &b)
{
    const int CTI = g_ct_inst(CT);
    const size_t lds = g_back_lds_bytes<MODE>(CTI, c->Kx + 1);
    if (CTI == 4) hipLaunchKernelGGL((k_gemm_back<MODE, 4, TT>), grid, dim3(512), lds, c->st, b);
    else if (CTI == 7) hipLaunchKernelGGL((k_gemm_back<MODE, 7, TT>), grid, dim3(512), lds, c->st, b);
    else if (CTI == 13) hipLaunchKernelGGL((k_gemm_back<MODE, 13, TT>), grid, dim3(512), lds, c->st, b);
    else hipLaunchKernelGGL((k_gemm_back<MODE, MB_CT, TT>), grid, dim3(512), lds, c->st, b);
}

// dynamic LDS limits of every back-projection instantiation (the largest
// exceed the 64 KB default)
template <int MODE, typename TT>
static bool g_back_attrs_t(int KX1)
{
    const void *fn[4] = {(const void *)k_gemm_back<MODE, 4, TT>, (const void *)k_gemm_back<MODE, 7, TT>,
                         (const void *)k_gemm_back<MODE, 13, TT>, (const void *)k_gemm_back<MODE, MB_CT, TT>};
    const int cti[4] = {4, 7, 13, MB_CT};
    for (int k = 0; k < 4; k++)
        if (hipFuncSetAttribute(fn[k], hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)g_back_lds_bytes<MODE>(cti[k], KX1)) != hipSuccess)
            return false;
    return true;
}

static bool g_back_attrs(mdff_ctx *)
{
    // the function attribute is process-wide: size it for the largest k box
    // the stored-operand path takes, (Kx+1) 6 2 <= 16 MB_CT
    const int KX1 = 16 * MB_CT / 12;
    return g_back_attrs_t<KM_F64, double>(KX1) && g_back_attrs_t<KM_F32, float>(KX1) &&
           g_back_attrs_t<KM_F16X3, float>(KX1);
}

// Block target of one stored-operand sum (mode, direction, tile count,
// ion set): on first use, candidates of 1/2 .. 3 blocks per compute unit
// are timed with HIP events (one warm launch, then three) and the fastest
// is kept.  `run` launches the sum (and, for the forward sum, its
// partial-sum pass) ungated; the launches only rewrite partial-sum buffers
// before the real launch writes them again.
template <typename F>
static int gtune_blocks(mdff_ctx *c, int mode, int dir, int tiles, int polar, F &&run)
{
    for (const auto &t : c->gtune)
        if (t.mode == mode && t.dir == dir && t.tiles == tiles && t.polar == polar) return t.blocks;
    const int cand[6] = {c->ncu / 2, 3 * c->ncu / 4, c->ncu, 3 * c->ncu / 2, 2 * c->ncu, 3 * c->ncu};
    int best = c->ncu;
    // time on a quiet device: real-space kernels still running on the second
    // stream would slow some candidates and not others
    if (c->st2) (void)hipStreamSynchronize(c->st2);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        // two sweeps over the candidates; each candidate keeps its best time
        float bt[6] = {1e30f, 1e30f, 1e30f, 1e30f, 1e30f, 1e30f};
        for (int sweep = 0; sweep < 2; sweep++) {
            int last = -1;
            for (int k = 0; k < 6; k++) {
                const int splits = run(cand[k], 0);   // the split count this target gives (0: no launch)
                if (splits == last) continue;         // same grid as the previous candidate
                last = splits;
                run(cand[k], 1);
                (void)hipEventRecord(e0, c->st);
                for (int r = 0; r < 3; r++) run(cand[k], 1);
                (void)hipEventRecord(e1, c->st);
                float ms = 0.0f;
                if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
                    bt[k] = std::min(bt[k], ms);
            }
        }
        float b = 1e30f;
        for (int k = 0; k < 6; k++)
            if (bt[k] < b) { b = bt[k]; best = cand[k]; }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    c->gtune.push_back({mode, dir, tiles, polar, best});
    return best;
}

// k_psum_reduce_cm over the FP32 partials of a deferred forward sum
static void launch_psum_cm(mdff_ctx *c, const KRedArgs &r, int C, int CH, int nch, const double *inv)
{
    const dim3 gp(cdiv(c->nkf, PSUM_PB));
    const int M = (c->Kx + 1) * 2 * C;
    const float *kp = (const float *)c->kpart;
#define PSUM_CM(CC, HH) hipLaunchKernelGGL((k_psum_reduce_cm<CC, HH>), gp, dim3(256), 0, c->st, r, kp, nch, M, inv, c->gcm)
    if (C == 1) { if (CH == 6) PSUM_CM(1, 6); else PSUM_CM(1, 3); }
    else if (C == 3) { if (CH == 6) PSUM_CM(3, 6); else PSUM_CM(3, 3); }
    else { if (CH == 6) PSUM_CM(4, 6); else PSUM_CM(4, 3); }
#undef PSUM_CM
}

// forward sum over stored operands; 0 when outside the kernel's shapes
static int launch_fwd_gemm(mdff_ctx *c, int polar, int lo, int C, const double *q_w, const double *mu_w,
                           const int *active, int *nchunk_out)
{
    const int M = (c->Kx + 1) * C * 2;
    const int MT = (M + 15) / 16;
    if (MT > MF_MT) return 0;
    const int mode = lo ? c->mf_lo : KM_F64;
    // stored YZ for the FP64 mode, and where a block's table slice does not
    // fit the staging area; the others form it from the tables
    const int CW = c->g_cw[mode == KM_F64 ? 0 : 1];
    const bool gen = mode != KM_F64 && g_fwd_gen_fits(CW, 2 * c->Kz + 1);
    if (!gen && ensure_yz(c, mode, polar, lo)) return 0;
    const int nsel = polar ? c->npol : c->n;
    const int nselp = c->nselp[polar];
    const size_t e = g_elt(mode);
    const size_t awb = (size_t)16 * MF_MT * 2 * std::max(c->nselp[0], c->nselp[1]) * e;
    if (galloc(c, &c->aw[mode][0], awb)) return 0;
    if (g_split_mode(mode) && galloc(c, &c->aw[mode][1], awb)) return 0;
    void **t = c->tab3[lo][polar];
    const int *idx = polar ? c->pidx : nullptr;
    dim3 ga(cdiv(nselp, 256), c->Kx + 1);
    const double *wsc = nullptr, *winv = nullptr;
    const int nwb0 = cdiv(nsel, 256);
    int nwb = (c->wpart_ready && c->wpart_nb > 0) ? c->wpart_nb : nwb0;
    if (mode == KM_F16X3) {
        // weights scaled into [-1, 1] per channel (maxima already in gwp
        // when the kernel that wrote the weights produced them)
        wsc = c->gwp;
        if (C == 1 && !polar && q_w == c->q) {
            // the charges alone (static field): their maxima never change
            if (!c->gwq_ok)
                hipLaunchKernelGGL(k_w_max, dim3(nwb0), dim3(256), 0, c->st, nsel, C, idx, q_w, mu_w, c->n, c->gwq,
                                   (const int *)nullptr);
            c->gwq_ok = 1;
            wsc = c->gwq;
            nwb = nwb0;
        } else if (!c->wpart_ready) {
            hipLaunchKernelGGL(k_w_max, dim3(nwb), dim3(256), 0, c->st, nsel, C, idx, q_w, mu_w, c->n, c->gwp,
                               active);
        }
        winv = c->gsc + 8;
    }
    if (mode == KM_F64)
        hipLaunchKernelGGL((k_aw_build<KM_F64, double>), ga, dim3(256), 0, c->st, nsel, nselp, C, (const double *)t[0],
                           idx, q_w, mu_w, c->n, c->aw[mode][0], c->aw[mode][1], wsc, nwb, c->gsc + 8, active);
    else if (mode == KM_F32)
        hipLaunchKernelGGL(k_aw_build16<KM_F32>, dim3(cdiv(nselp, 256 * 2), c->Kx + 1, C), dim3(256), 0, c->st, nsel,
                           nselp, C, (const float *)t[0], idx, q_w, mu_w, c->n, c->aw[mode][0], c->aw[mode][1], wsc,
                           nwb, c->gsc + 8, active);
    else
        hipLaunchKernelGGL(k_aw_build16<KM_F16X3>, dim3(cdiv(nselp, 256 * 4), c->Kx + 1, C), dim3(256), 0, c->st, nsel,
                           nselp, C, (const float *)t[0], idx, q_w, mu_w, c->n, c->aw[mode][0], c->aw[mode][1], wsc,
                           nwb, c->gsc + 8, active);
    GFwdArgs a{};
    a.aw = c->aw[mode][0]; a.aw_lo = c->aw[mode][1];
    a.yz = c->yz[mode][polar][0]; a.yz_lo = c->yz[mode][polar][1];
    a.gty = (const float *)t[1]; a.gtz = (const float *)t[2];
    a.nsel = nsel; a.NY = 2 * c->Ky + 1; a.NZ = 2 * c->Kz + 1;
    a.MT = MT; a.M = M; a.P = c->P; a.C = C; a.ldk = 2 * nselp;
    a.part = c->kpart; a.active = active;
    const int gx = (c->P + 64 * CW - 1) / (64 * CW);
    const size_t nval = (size_t)M * c->P;
    // K split to reach `blocks` blocks (K ranges of an even number of steps:
    // the kernel alternates two register sets; ldk = 2 nselp is a multiple
    // of 2 GK), then the chunk sums; with go = 0 only the split count
    // defer: leave the FP32 partials to k_psum_reduce_cm (the column maxima
    // it fills are zeroed by the sum's first block)
    const bool defer = c->psum_defer && mode == KM_F16X3;
    // three dipole channels: the sum writes rho = i k.S_mu (one channel)
    const bool comb = defer && C == 3 && !c->no_fuse;
    a.comb = comb;
    a.kvec = c->kvec;
    a.cinv = winv;
    // a deferred forward's block count is timed with the reduction that sums
    // its partials in the call (k_psum_reduce_cm; its outputs are written
    // again by the call's own reduction)
    KRedArgs tr{};
    tr.Kx = c->Kx; tr.Ky = c->Ky; tr.Kz = c->Kz; tr.P = c->P; tr.nchunk = 1; tr.C = C;
    tr.part = c->ksum; tr.kvec = c->kvec; tr.Ak = c->Ak; tr.kcoe = c->kcoe; tr.W = c->W; tr.ek = c->ek;
    tr.do_stress = 0; tr.active = nullptr;
    auto run = [&](int blocks, int go, const int *act, bool probe, bool psum = true) {
        int nch = std::max(1, std::min(64, (blocks + gx - 1) / gx));
        GFwdArgs ga2 = a;
        ga2.active = act;
        if (!psum) { ga2.cmz = c->gcm; ga2.ncmz = 6 * (c->Kx + 1); }
        ga2.kchunk = ((ga2.ldk + nch - 1) / nch + 2 * GK - 1) / (2 * GK) * (2 * GK);
        nch = (ga2.ldk + ga2.kchunk - 1) / ga2.kchunk;
        if (!go) return nch;
        dim3 grid(gx, nch);
        const int reps = probe_reps(c, probe);
        if (probe) probe_mark(c, 0);
        for (int r = 0; r < reps; r++) {
            if (mode == KM_F64) gfwd_cw<KM_F64>(c, grid, CW, MT, ga2);
            else if (mode == KM_F32) { if (gen) gfwd_cw<KM_F32, true>(c, grid, CW, MT, ga2); else gfwd_cw<KM_F32>(c, grid, CW, MT, ga2); }
            else { if (gen) gfwd_cw<KM_F16X3, true>(c, grid, CW, MT, ga2); else gfwd_cw<KM_F16X3>(c, grid, CW, MT, ga2); }
        }
        if (probe) probe_mark(c, 1);
        if (!psum) return nch;
        if (defer) {
            if (comb) {
                KRedArgs t1 = tr;
                t1.C = 1;
                launch_psum_cm(c, t1, 1, c->psum_ch, nch, nullptr);
            } else {
                launch_psum_cm(c, tr, C, c->psum_ch, nch, winv);
            }
            return nch;
        }
        if (mode == KM_F64)
            hipLaunchKernelGGL(k_part_sum<double>, dim3(cdiv(nval, 256)), dim3(256), 0, c->st,
                               (const double *)c->kpart, nch, M, c->P, c->ksum, C, winv, act);
        else
            hipLaunchKernelGGL(k_part_sum<float>, dim3(cdiv(nval, 256)), dim3(256), 0, c->st, (const float *)c->kpart,
                               nch, M, c->P, c->ksum, C, winv, act, c->gcm, 6 * (c->Kx + 1));
        return nch;
    };
    const int fb = c->g_fblocks > 0 ? c->g_fblocks
                                    : gtune_blocks(c, mode, 0, MT, polar, [&](int b, int go) {
                                          return run(b, go, nullptr, false);
                                      });
    const int nch = run(fb, 1, active, true, !defer);
    c->fwd_part = defer ? 2 : 1;
    c->fwd_nch = nch;
    c->fwd_inv = winv;
    c->fwd_comb = comb;
    *nchunk_out = 1;
    return 1;
}

// WM columns of a mode (both parts for the split modes)
static int wm_alloc(mdff_ctx *c, int mode)
{
    const size_t wmb = (size_t)16 * MB_CT * 2 * c->Pk * g_elt(mode);
    int rc;
    if ((rc = galloc(c, &c->wm[mode][0], wmb)) != MDFF_OK) return rc;
    if (g_split_mode(mode) && (rc = galloc(c, &c->wm[mode][1], wmb)) != MDFF_OK) return rc;
    return MDFF_OK;
}

// back-projection over stored operands (WM columns formed here unless
// wm_built: k_sk_reduce_wm wrote them); 0 when outside the kernel's shapes
static int launch_back_gemm(mdff_ctx *c, int polar, int lo, int NCH, const int *active, int *nsplit_out,
                            bool wm_built, T2Rows *t2r)
{
    const int CH = NCH == 9 ? 6 : 3;
    const int NC = (c->Kx + 1) * CH * 2;
    const int CT = (NC + 15) / 16;
    if (CT > MB_CT || !g_back_fits(c, CH)) return 0;
    const int mode = lo ? c->mf_lo : KM_F64;
    // stored YZt in the FP64 mode only (the others form it from the tables)
    if (mode == KM_F64 && ensure_yz(c, mode, polar, lo)) return 0;
    const int nsel = polar ? c->npol : c->n;
    const int nselp = c->nselp[polar];
    if (wm_alloc(c, mode) != MDFF_OK) return 0;
    dim3 gw(cdiv(c->Pk, 256), (c->Kx + 1) * CH);
    if (wm_built) {
        // columns written by k_sk_reduce_wm
    } else if (mode == KM_F16X3)
        hipLaunchKernelGGL(k_wm_build_scaled<KM_F16X3>, dim3((c->Kx + 1) * CH), dim3(256), 0, c->st, c->P, c->Pk,
                           c->Ky, c->Kz, CH, c->W, c->wm[mode][0], c->wm[mode][1], c->gsc + 16, active);
    else if (mode == KM_F64)
        hipLaunchKernelGGL(k_wm_build<KM_F64>, gw, dim3(256), 0, c->st, c->P, c->Pk, c->Ky, c->Kz, CH, c->W,
                           c->wm[mode][0], c->wm[mode][1], active);
    else
        hipLaunchKernelGGL(k_wm_build<KM_F32>, gw, dim3(256), 0, c->st, c->P, c->Pk, c->Ky, c->Kz, CH, c->W,
                           c->wm[mode][0], c->wm[mode][1], active);
    GBackArgs b{};
    b.yzt = c->yzt[mode][polar][0]; b.yzt_lo = c->yzt[mode][polar][1];
    b.gty = (const float *)c->tab3[lo][polar][1]; b.gtz = (const float *)c->tab3[lo][polar][2];
    b.NY = 2 * c->Ky + 1; b.NZ = 2 * c->Kz + 1;
    b.wm = c->wm[mode][0]; b.wm_lo = c->wm[mode][1];
    b.tx = c->tab3[lo][polar][0];
    b.csc = mode == KM_F16X3 ? c->gsc + 16 : nullptr;
    b.nsel = nsel; b.CH = CH; b.NC = NC; b.CT = CT; b.ldk = 2 * c->Pk;
    b.part = c->bpart; b.active = active;
    const int gx = nselp / 128;
    // pair-tensor rows: 32 polarisable ions per block over gx blocks a row
    const bool t2rows = t2r && polar && c->t2_ok;
    const int nt2 = t2rows ? cdiv(c->npol, 32 * gx) : 0;
    if (t2rows) {
        b.t2_npol = c->npol; b.t2_capP = c->capP; b.t2_n = c->n;
        b.t2 = t2r->t2; b.t2_cnt = c->cntp; b.t2_idx = c->idxp; b.t2_pidx = c->pidx;
        b.t2mu = t2r->mu; b.t2out = t2r->out;
    }
    // K split to reach `blocks` blocks (even step counts as in the forward
    // sum; ldk = 2 Pk is a multiple of 2 GK); with go = 0 only the count
    auto run = [&](int blocks, int go, const int *act, bool probe) {
        int nsp = std::max(1, std::min(c->nsplit_max, (blocks + gx - 1) / gx));
        GBackArgs gb2 = b;
        gb2.active = act;
        gb2.kchunk = ((gb2.ldk + nsp - 1) / nsp + 2 * GK - 1) / (2 * GK) * (2 * GK);
        nsp = (gb2.ldk + gb2.kchunk - 1) / gb2.kchunk;
        if (!go) return nsp;
        gb2.t2_y0 = nsp;
        dim3 grid(gx, nsp + nt2);
        const int reps = probe_reps(c, probe);
        if (probe) probe_mark(c, 0);
        for (int r = 0; r < reps; r++) {
            if (mode == KM_F64) gback_ct<KM_F64, double>(c, grid, CT, gb2);
            else if (mode == KM_F32) gback_ct<KM_F32, float>(c, grid, CT, gb2);
            else gback_ct<KM_F16X3, float>(c, grid, CT, gb2);
        }
        if (probe) probe_mark(c, 1);
        return nsp;
    };
    const int bb = c->g_bblocks > 0 ? c->g_bblocks
                                    : gtune_blocks(c, mode, 1, CT, polar, [&](int bl, int go) {
                                          return run(bl, go, nullptr, false);
                                      });
    const int ns = run(bb, 1, active, true);
    *nsplit_out = ns;
    if (t2rows) t2r->done = 1;
    return 1;
}

// reciprocal part of multipole_ES (ewald.c:876-1051) on the selected ions.
// polar: use the polarisable-ion tables (SCF); C: channels of S(k);
// NCH: 3 (field) or 9 (field + gradient); do_ef/do_f outputs; efg optional.
// total field formed in the back-projection's last kernel (self + dir +
// rec + base), instead of a separate k_field_sum
struct FieldFuse {
    double *ef;
    const double *mu, *base;
    double *pq_part;   // optional: the CG step's p.q partial sums (k_sk_back_fin)
    const void *t2;    // optional: pair tensors whose field of mu the finish forms (k_sk_back_fin_t2)
    int t2f32;         // t2 in FP32
    int join;          // join the second stream (its real-space field in ef_dir) before the finish
};

static int rec_part(mdff_ctx *c, int polar, int lo, int C, int NCH, const double *q_w, const double *mu_w,
                    const double *q_f, const double *mu_f, int do_ef, int do_f, int do_stress, double *efg_out,
                    const int *active, const FieldFuse *ff = nullptr)
{
    int npq = 0;
    int nsel = polar ? c->npol : c->n;
    int chunk = polar ? c->chunk_p : c->chunk;
    int nchunk = polar ? c->nchunk_p : c->nchunk;
    SkArgs a;
    a.nsel = nsel;
    a.Kx = c->Kx; a.Ky = c->Ky; a.Kz = c->Kz; a.P = c->P;
    a.tx = c->tab3[lo][polar][0];
    a.ty = c->tab3[lo][polar][1];
    a.tz = c->tab3[lo][polar][2];
    a.idx = polar ? c->pidx : nullptr;
    a.q = q_w; a.mu = mu_w; a.n = c->n;
    a.chunk = chunk;
    a.part = c->kpart;
    a.active = active;
    bool mixed = lo != 0;
    const int bmode = lo ? c->mf_lo : KM_F64;
    const int CHb = NCH == 9 ? 6 : 3;
    c->fwd_part = 0;
    // split FP16 with a back-projection: the reduction sums the forward's
    // partials itself (k_psum_reduce_cm)
    const bool cm_path = (do_ef || do_f) && c->use_mfma && bmode == KM_F16X3 &&
                         g_back_fits(c, CHb) && wm_alloc(c, bmode) == MDFF_OK;
    c->psum_defer = cm_path && !c->no_fuse && (C == 1 || C == 3 || C == 4);
    c->psum_ch = CHb;
    const bool fwd_mf = launch_fwd_mfma(c, polar, lo, C, q_w, mu_w, active, &nchunk);
    c->psum_defer = false;
    c->wpart_ready = 0;
    c->wpart_nb = 0;
    if (!fwd_mf) {
        if (C == 1) { if (mixed) launch_fwd_t<float, 1>(c, a, nchunk); else launch_fwd_t<double, 1>(c, a, nchunk); }
        else if (C == 3) { if (mixed) launch_fwd_t<float, 3>(c, a, nchunk); else launch_fwd_t<double, 3>(c, a, nchunk); }
        else { if (mixed) launch_fwd_t<float, 4>(c, a, nchunk); else launch_fwd_t<double, 4>(c, a, nchunk); }
    }
    KRedArgs r;
    r.Kx = c->Kx; r.Ky = c->Ky; r.Kz = c->Kz; r.P = c->P; r.nchunk = nchunk; r.C = C;
    const bool gemm_fwd = c->fwd_part != 0;
    const bool raw = c->fwd_part == 2;
    r.part = gemm_fwd ? c->ksum : c->kpart; r.kvec = c->kvec;
    c->fwd_part = 0; r.Ak = c->Ak; r.kcoe = c->kcoe; r.W = c->W; r.ek = c->ek;
    r.do_stress = do_stress; r.active = active;
    // the back-projection's WM columns come out of the reduction: for split
    // FP16 (column-scaled) the reduction gives the column maxima and
    // k_wm_scaled the columns (the reduction sums the forward's partials
    // itself when the forward left them, raw); the other modes form them in
    // the reduction (k_sk_reduce_wm)
    bool wm_built = false;
    if (cm_path && gemm_fwd) {
        void *w0 = c->wm[bmode][0], *w1 = c->wm[bmode][1];
        const dim3 gr(cdiv(c->nkf, 256));
        if (raw && c->fwd_comb) {
            KRedArgs r1 = r;
            r1.C = 1;
            launch_psum_cm(c, r1, 1, CHb, c->fwd_nch, nullptr);
        } else if (raw) {
            launch_psum_cm(c, r, C, CHb, c->fwd_nch, c->fwd_inv);
        } else if (CHb == 6) {
            hipLaunchKernelGGL(k_sk_reduce_cm<6>, gr, dim3(256), 0, c->st, r, c->gcm);
        } else {
            hipLaunchKernelGGL(k_sk_reduce_cm<3>, gr, dim3(256), 0, c->st, r, c->gcm);
        }
        hipLaunchKernelGGL(k_wm_scaled, dim3(cdiv(c->Pk, 256), (c->Kx + 1) * CHb), dim3(256), 0, c->st, c->P, c->Pk,
                           c->Ky, c->Kz, CHb, c->W, c->gcm, w0, w1, c->gsc + 16, active);
        wm_built = true;
    } else if ((do_ef || do_f) && gemm_fwd && c->use_mfma && bmode != KM_F16X3 &&
        g_back_fits(c, CHb) && wm_alloc(c, bmode) == MDFF_OK) {
        void *w0 = c->wm[bmode][0], *w1 = c->wm[bmode][1];
        const dim3 gr(cdiv(c->nkf, 256));
        if (bmode == KM_F64) hipLaunchKernelGGL(k_sk_reduce_wm<KM_F64>, gr, dim3(256), 0, c->st, r, CHb, c->Pk, w0, w1);
        else hipLaunchKernelGGL(k_sk_reduce_wm<KM_F32>, gr, dim3(256), 0, c->st, r, CHb, c->Pk, w0, w1);
        wm_built = true;
    } else {
        hipLaunchKernelGGL(k_sk_reduce, dim3(cdiv(c->nkf, 256)), dim3(256), 0, c->st, r);
    }
    if (do_ef || do_f) {
        BackArgs b;
        b.nsel = nsel; b.Kx = c->Kx; b.Ky = c->Ky; b.Kz = c->Kz; b.P = c->P;
        b.tx = a.tx; b.ty = a.ty; b.tz = a.tz;
        b.W = c->W;
        b.rows_per_split = c->rows_per_split;
        b.part = c->bpart;
        b.active = active;
        int nsplit = c->nsplit;
        // the FP32 pair-tensor field of the finish in extra rows of the
        // back-projection's launch, when it takes them
        T2Rows t2r{ff && ff->t2 && ff->t2f32 && !c->no_fuse ? (const float *)ff->t2 : nullptr, ff ? ff->mu : nullptr,
                   c->t2dir, 0};
        if (!launch_back_mfma(c, polar, lo, NCH, active, &nsplit, wm_built, t2r.t2 ? &t2r : nullptr)) {
            if (NCH == 9) { if (mixed) launch_back_t<float, 9>(c, b, nsel); else launch_back_t<double, 9>(c, b, nsel); }
            else { if (mixed) launch_back_t<float, 3>(c, b, nsel); else launch_back_t<double, 3>(c, b, nsel); }
        }
        BackFinArgs fa;
        fa.n = c->n; fa.nsel = nsel; fa.nsplit = nsplit; fa.NCH = NCH;
        fa.idx = a.idx; fa.c = c->cd; fa.tfpiV = c->tfpiV; fa.part = c->bpart;
        fa.q = q_f; fa.mu = mu_f;
        fa.ef = c->ef_rec; fa.efg = efg_out; fa.f = c->f_rec;
        fa.do_ef = do_ef; fa.do_f = do_f; fa.active = active;
        fa.ef_total = ff ? ff->ef : nullptr;
        fa.mu_self = ff ? ff->mu : nullptr;
        fa.ef_dir = c->ef_dir;
        fa.base = ff ? ff->base : nullptr;
        fa.selfa2 = c->selfa2;
        fa.type = c->type;
        fa.T = c->tab_d;
        fa.pq_part = ff ? ff->pq_part : nullptr;
        if (ff && ff->join) aux_join(c);
        if (t2r.done) {
            // the pair-tensor field came out of the back-projection's launch
            fa.ef_dir = c->t2dir;
            hipLaunchKernelGGL(k_sk_back_fin, dim3(cdiv(nsel, 256)), dim3(256), 0, c->st, fa);
            if (fa.pq_part) npq = cdiv(nsel, 256);
        } else if (ff && ff->t2) {
            const int L = fin_t2_lanes(c);
            if (fa.pq_part) npq = cdiv(nsel * L, 256);
            const dim3 gf(cdiv((size_t)nsel * L, 256));
            if (ff->t2f32) {
                const float *t2 = (const float *)ff->t2;
                if (L == 8) hipLaunchKernelGGL((k_sk_back_fin_t2<float, 8>), gf, dim3(256), 0, c->st, fa, c->capP, c->npol, c->cntp, c->idxp, t2);
                else if (L == 64) hipLaunchKernelGGL((k_sk_back_fin_t2<float, 64>), gf, dim3(256), 0, c->st, fa, c->capP, c->npol, c->cntp, c->idxp, t2);
                else if (L == 32) hipLaunchKernelGGL((k_sk_back_fin_t2<float, 32>), gf, dim3(256), 0, c->st, fa, c->capP, c->npol, c->cntp, c->idxp, t2);
                else hipLaunchKernelGGL((k_sk_back_fin_t2<float, 16>), gf, dim3(256), 0, c->st, fa, c->capP, c->npol, c->cntp, c->idxp, t2);
            } else {
                const double *t2 = (const double *)ff->t2;
                if (L == 8) hipLaunchKernelGGL((k_sk_back_fin_t2<double, 8>), gf, dim3(256), 0, c->st, fa, c->capP, c->npol, c->cntp, c->idxp, t2);
                else if (L == 64) hipLaunchKernelGGL((k_sk_back_fin_t2<double, 64>), gf, dim3(256), 0, c->st, fa, c->capP, c->npol, c->cntp, c->idxp, t2);
                else if (L == 32) hipLaunchKernelGGL((k_sk_back_fin_t2<double, 32>), gf, dim3(256), 0, c->st, fa, c->capP, c->npol, c->cntp, c->idxp, t2);
                else hipLaunchKernelGGL((k_sk_back_fin_t2<double, 16>), gf, dim3(256), 0, c->st, fa, c->capP, c->npol, c->cntp, c->idxp, t2);
            }
        } else {
            hipLaunchKernelGGL(k_sk_back_fin, dim3(cdiv(nsel, 256)), dim3(256), 0, c->st, fa);
            if (fa.pq_part) npq = cdiv(nsel, 256);
        }
    }
    return npq;
}

static void short_range(mdff_ctx *c)
{
    int n = c->n;
    if (!c->lnonbonded) return;
    ShortArgs a;
    a.n = n; a.cap = c->cap; a.c = c->cd; a.s = c->s; a.type = c->type; a.nbr = c->nbr; a.cnt = c->cnt;
    a.T = c->tab_d; a.rc2 = c->srcutsq; a.f = c->f_sr; a.acc = c->acc_sr;
    int L = c->lanes;
    unsigned blocks = cdiv((size_t)n * L, 256);
    bool mixed = lo_of(c);
    if (L == 8) {
        if (mixed) hipLaunchKernelGGL((k_short<float, 8>), dim3(blocks), dim3(256), 0, c->st, a);
        else hipLaunchKernelGGL((k_short<double, 8>), dim3(blocks), dim3(256), 0, c->st, a);
    } else if (L == 16) {
        if (mixed) hipLaunchKernelGGL((k_short<float, 16>), dim3(blocks), dim3(256), 0, c->st, a);
        else hipLaunchKernelGGL((k_short<double, 16>), dim3(blocks), dim3(256), 0, c->st, a);
    } else if (L == 64) {
        if (mixed) hipLaunchKernelGGL((k_short<float, 64>), dim3(blocks), dim3(256), 0, c->st, a);
        else hipLaunchKernelGGL((k_short<double, 64>), dim3(blocks), dim3(256), 0, c->st, a);
    } else {
        if (mixed) hipLaunchKernelGGL((k_short<float, 32>), dim3(blocks), dim3(256), 0, c->st, a);
        else hipLaunchKernelGGL((k_short<double, 32>), dim3(blocks), dim3(256), 0, c->st, a);
    }
    colsum(c, c->acc_sr, n, n, 14, c->scal + S_U_BH);
}

// static field of momentpolaSCF pim.c:216-218: charges (+ static dipoles)
// with damping, field only
static void static_field(mdff_ctx *c, int lo)
{
    // real space on the second stream, k space here
    aux_fork(c);
    {
        OnAux g(c);
        DirArgs a = dir_args(c, c->q, c->mu_static, c->ldmp, 0, nullptr);
        launch_dir(c, a, lo, 1, 0, 0, 0);
    }
    int C = c->ldip_static ? 4 : 1;
    // the back-projection's finish forms ef_stat = self + dir + rec (after
    // joining the real-space field), in place of k_field_sum
    FieldFuse ff{c->ef_stat, c->mu_static, nullptr, nullptr, nullptr, 0, 1};
    rec_part(c, 0, lo, C, 3, c->q, c->mu_static, c->q, c->mu_static, 1, 0, 0, nullptr, nullptr, &ff);
}

// ASPC order and coefficients of extrapolate_dipole_aspc pim.c:324-379
static int aspc_args(mdff_ctx *c, int istep, AspcArgs *asp)
{
    int order = c->p.extrapolate_order;
    int eo = (istep >= order + 1) ? order : (istep ? istep - 1 : 0);
    double B[6] = {0, 0, 0, 0, 0, 0}, Wc = 0.0;
    switch (eo) {
    case 0: B[0] = 1.0; Wc = 0.0; break;
    case 1: B[0] = 2.0; B[1] = -1.0; Wc = 2.0 / 3.0; break;
    case 2: B[0] = 2.5; B[1] = -2.0; B[2] = 0.5; Wc = 0.6; break;
    case 3: B[0] = 2.8; B[1] = -2.8; B[2] = 1.2; B[3] = -0.2; Wc = 4.0 / 7.0; break;
    case 4: B[0] = 3.0; B[1] = -24.0 / 7.0; B[2] = 27.0 / 14.0; B[3] = -4.0 / 7.0; B[4] = 1.0 / 14.0; Wc = 5.0 / 9.0; break;
    case 5:
        B[0] = 22.0 / 7.0; B[1] = -55.0 / 14.0; B[2] = 55.0 / 21.0; B[3] = -22.0 / 21.0; B[4] = 5.0 / 21.0;
        B[5] = -1.0 / 42.0; Wc = 6.0 / 11.0;
        break;
    default: return MDFF_ERR_CONFIG;
    }
    asp->n = c->n;
    asp->eo = eo;
    asp->shift = (eo != order) ? eo + 1 : eo;
    asp->W = Wc;
    for (int k = 0; k < 6; k++) asp->B[k] = B[k];
    return MDFF_OK;
}

// alphaES of the inner real-space sweeps of scfkO (pim.c:693-694: the value
// computed from epsw just before is overwritten by this constant)
#define KO_ALPHA 0.001

// dipole tensors of all polarisable pairs at the current positions (per
// force call, after k_frac), and for scfkO the same at KO_ALPHA; rows hold
// the list capacity, so they never overflow
// whether the mixed form's SCF runs the anchored CG (solver 2 with
// symmetric positive definite polarisabilities), and whether the SCF
// evaluates the induced field in FP64: the FP64 form, scfkO (FP64 in both
// forms), and the mixed form but anchored CG with a low-precision anchor
static bool scf_anchored_ok(const mdff_ctx *c) { return c->scf_solver == 2 && c->cg_ok && c->npol > 0; }

static bool scf_uses_fp64(const mdff_ctx *c)
{
    return !lo_of(c) || c->p.algo_pim == MDFF_PIM_SCFKO || !(scf_anchored_ok(c) && c->mf_anchor);
}

// T2 pair tensors of the polarisable ions; the FP64 copy only when the SCF
// (or fp64, for kernel timing) needs it
static int build_t2(mdff_ctx *c, bool fp64 = false)
{
    c->t2_ok = 0;
    if (c->npol == 0) return MDFF_OK;
    const bool ko = c->p.algo_pim == MDFF_PIM_SCFKO;
    if (c->capP != c->cap || (ko && !c->t2ko)) {
        dfree(c->idxp);
        dfree(c->t2);
        dfree(c->t2ko);
        if (c->t2f) { (void)hipFree(c->t2f); c->t2f = nullptr; }
        size_t ne = (size_t)c->npol * c->cap;
        if (hipMalloc((void **)&c->idxp, sizeof(int) * ne) != hipSuccess ||
            hipMalloc((void **)&c->t2, sizeof(double) * 6 * ne) != hipSuccess)
            return MDFF_ERR_NOMEM;
        if (lo_of(c) && hipMalloc((void **)&c->t2f, sizeof(float) * 6 * ne) != hipSuccess) return MDFF_ERR_NOMEM;
        if (ko && hipMalloc((void **)&c->t2ko, sizeof(double) * 6 * ne) != hipSuccess) return MDFF_ERR_NOMEM;
        c->capP = c->cap;
    }
    T2Args a;
    a.n = c->n; a.cap = c->cap; a.capP = c->capP; a.c = c->cd; a.s = c->s; a.nbr = c->nbr; a.cnt = c->cnt;
    a.pidx = c->pidx; a.pinv = c->pinv; a.alpha = c->alpha; a.rc2 = c->lrcutsq; a.npol = c->npol;
    a.cntp = c->cntp; a.idxp = c->idxp; a.t2f = lo_of(c) ? c->t2f : nullptr; a.overflow = c->t2_ovf;
    a.t2 = (fp64 || !a.t2f || scf_uses_fp64(c)) ? c->t2 : nullptr;
    // FP32 arithmetic when only the FP32 tensors are built
    if (a.t2) hipLaunchKernelGGL(k_t2_build<double>, dim3(cdiv(c->npol, 4)), dim3(256), 0, c->st, a);
    else hipLaunchKernelGGL(k_t2_build<float>, dim3(cdiv(c->npol, 4)), dim3(256), 0, c->st, a);
    if (ko) {
        // the same pairs at the alpha of induced_moment_inner_kO (pim.c:693-694);
        // the pair lists come out identical
        a.alpha = KO_ALPHA;
        a.t2 = c->t2ko;
        a.t2f = nullptr;
        hipLaunchKernelGGL(k_t2_build<double>, dim3(cdiv(c->npol, 4)), dim3(256), 0, c->st, a);
    }
    c->t2_ok = 1;
    return MDFF_OK;
}

// induced field of dipoles mu (no charges, no damping, pim.c:258-260) at
// the polarisable ions: ef = base + self + dir + rec
// with pq_part, the finish also writes the partial sums of
// mu . (pol^-T mu - ef); returns how many (0 when none were written)
static int induced_field(mdff_ctx *c, int lo, const double *mu, const double *base, double *ef, const int *act,
                         double *pq_part = nullptr)
{
    int n = c->n;
    if (c->t2_ok) {
        // the total field at the polarisable ions comes out of the back-
        // projection's finish, with the real-space field of mu from the
        // pair tensors (FP32 tensors: extra rows of the back-projection's
        // launch; else formed in the finish)
        FieldFuse ff{ef, mu, base, pq_part, lo ? (const void *)c->t2f : (const void *)c->t2, lo ? 1 : 0};
        return rec_part(c, 1, lo, 3, 3, c->qz, mu, c->qz, mu, 1, 0, 0, nullptr, act, &ff);
    }
    DirArgs a = dir_args(c, c->qz, mu, 0, 1, act);
    launch_dir(c, a, lo, 1, 0, 0, 0);
    rec_part(c, 1, lo, 3, 3, c->qz, mu, c->qz, mu, 1, 0, 0, nullptr, act);
    hipLaunchKernelGGL(k_field_sum, dim3(cdiv(3 * n, 256)), dim3(256), 0, c->st, n, c->selfa2, mu, c->ef_dir,
                       c->ef_rec, base, ef, act);
    return 0;
}

// rmsd of mu_ind against ef (and u_pol when upol is given) and the loop
// condition of momentpolaSCF pim.c:236, in one launch (k_scf_check)
static void scf_check(mdff_ctx *c, const double *upol, int maxit, int minit)
{
    const int nb = cdiv(c->n, 256);
    hipLaunchKernelGGL(k_scf_check, dim3(nb), dim3(256), 0, c->st, c->n, c->type, c->tab_d, c->mu_ind, c->ef, upol,
                       c->ctl, c->chk_part, c->chk_cnt, c->scal, c->npol, maxit, minit, c->p.conv_tol_ind);
}

static void scf_finish(mdff_ctx *c, bool store_hist = true)
{
    int n = c->n;
    // dipia_ind[ia][0] = mu_ind (pim.c:281-285)
    if (store_hist)
        hipLaunchKernelGGL(k_store_hist, dim3(cdiv(3 * n, 256)), dim3(256), 0, c->st, n, c->mu_ind, c->hist);
    c->last_nscf = c->ctl_h[C_ISCF];
    c->last_rmsd = sqrt(c->scal_h[S_RMSD2] / (double)(c->npol > 0 ? c->npol : 1));
}

// the SCF loop of momentpolaSCF pim.c:236-278 on the device (FP64 form,
// the reference's Jacobi schedule), gated by the convergence flag.  The
// host first launches as many iterations as the last SCF took (consecutive
// MD steps need about the same number), then polls the flag every
// scf_chunk iterations; iterations past convergence are skipped on the
// device, so the schedule never changes the result.
static int scf_loop(mdff_ctx *c, int istep)
{
    int n = c->n;
    const int *act = c->ctl + C_SCF_ACTIVE;
    int maxit = c->p.max_scf_pim_iter, minit = c->p.min_scf_pim_iter;
    hipLaunchKernelGGL(k_scf_begin, dim3(1), dim3(1), 0, c->st, c->ctl, (maxit > 0 || minit > 0) ? 1 : 0);
    AspcArgs asp;
    if (aspc_args(c, istep, &asp)) return MDFF_ERR_CONFIG;
    int launched = 0;
    int limit = std::max(maxit, minit) + 1;
    while (launched < limit) {
        int chunk = launched == 0 ? std::max(std::max(1, std::min(minit, c->scf_chunk)), c->hint_fp64)
                                  : c->scf_chunk;
        for (int it = 0; it < chunk && launched < limit; it++, launched++) {
            int first = launched == 0;
            hipLaunchKernelGGL(k_scf_mu, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, first, asp, c->type, c->tab_d,
                               c->hist, c->ef, c->mu_ind, c->upol_i, act);
            induced_field(c, 0, c->mu_ind, c->ef_stat, c->ef, act);
            if (first)
                hipLaunchKernelGGL(k_scf_correct, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, asp, c->type,
                                   c->tab_d, c->ef, c->mu_ind, act);
            // rmsd, u_pol of this iteration's dipoles and the loop condition
            scf_check(c, c->upol_i, maxit, minit);
        }
        int rc = sync_scal(c);
        if (rc) return rc;
        if (!c->ctl_h[C_SCF_ACTIVE]) break;
    }
    scf_finish(c);
    c->n_scf_fp64 += c->last_nscf;
    c->hint_fp64 = c->last_nscf;
    if (c->last_rmsd > c->p.conv_tol_ind) return MDFF_ERR_SCF;
    return MDFF_OK;
}

// Mixed form, anchored conjugate gradients (the default): one evaluation of
// the field of the ASPC predictor (pim.c:382-400), the anchor, in FP64 or
// (mf_anchor, the default) in a matrix-core mode over the FP32 tables and
// pair tensors, then conjugate-gradient steps on (pol^-T - T) mu = E_stat
// preconditioned by pol^T, each applying the low-precision T to the search
// direction only and carrying the field along (E <- E + gamma T p).  The
// rmsd of get_rmsd_mu (pim.c:431-449) is formed from the updated dipoles and
// the carried field after every step and the loop condition of pim.c:236
// applied to it on the device, one SCF iteration per step.  That rmsd is the
// residual of the low-precision system: the static field (FP32 real space,
// split-FP16 k space), the anchor and the steps each carry their own
// rounding, of order 1e-7 of the field, so the true FP64 residual of the
// final dipoles is larger than the tolerance asks; mdff_set_scf_check has it
// evaluated in FP64 (thermo rmsd_fp64).  The host launches as many steps as
// the last SCF took before it polls.
// k_sk_back_fin forms the CG step's p.q partial sums in its blocks of 256
// polarisable ions, the blocks of the CG kernels
static_assert(CG_BLOCK == 256, "CG partial-sum blocks");

static int scf_loop_anchored(mdff_ctx *c, int istep, bool refine = false)
{
    int n = c->n, np = c->npol;
    int *ctl = c->ctl;
    const int *act = c->ctl + C_SCF_ACTIVE;
    int maxit = c->p.max_scf_pim_iter, minit = c->p.min_scf_pim_iter;
    // the FP64 form runs the same solver with FP64 anchor and steps
    const int lo = lo_of(c);
    const int anchor = (lo && !refine) ? c->mf_anchor : 0;
    if (refine) {
        // FP64 refinement: the converged dipoles are the start, the anchor
        // is their FP64 field, and the loop may stop at once
        minit = 0;
        hipLaunchKernelGGL(k_scf_begin, dim3(1), dim3(1), 0, c->st, ctl, 1);
    } else {
        AspcArgs asp;
        if (aspc_args(c, istep, &asp)) return MDFF_ERR_CONFIG;
        // the predictor opens the SCF (k_scf_begin) when the loop runs at
        // all; a split-FP16 anchor scales its weights by the predictor's
        // channel maxima, which the predictor writes (all ions; zero where
        // not polarisable)
        const bool open = maxit > 0 || minit > 0;
        if (!open) hipLaunchKernelGGL(k_scf_begin, dim3(1), dim3(1), 0, c->st, ctl, 0);
        const bool wmu = anchor == KM_F16X3;
        hipLaunchKernelGGL(k_scf_mu, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, 1, asp, c->type, c->tab_d, c->hist,
                           c->ef, c->mu_ind, c->upol_i, act, wmu ? c->gwp : nullptr, open ? ctl : nullptr);
    }
    // the anchor: the field of the predicted dipoles in FP64, or with
    // mf_anchor over the FP32 tables in that matrix-core mode
    if (anchor) {
        const int keep = c->mf_lo;
        const bool wmu = anchor == KM_F16X3;
        c->mf_lo = anchor;
        c->wpart_ready = wmu;
        c->wpart_nb = wmu ? cdiv(n, 256) : 0;
        induced_field(c, 1, c->mu_ind, c->ef_stat, c->ef, act);
        c->mf_lo = keep;
    } else {
        induced_field(c, 0, c->mu_ind, c->ef_stat, c->ef, act);
    }
    scf_check(c, nullptr, maxit, minit);
    const int nb = cdiv(np, CG_BLOCK);
    // split-FP16 steps: the kernels that write p also give the channel
    // maxima of p that scale the forward sum's weights
    double *wp = (lo && c->mf_inner == KM_F16X3) ? c->gwp : nullptr;
    double *rz[2] = {c->cg_part, c->cg_part + nb}, *rm = c->cg_part + 3 * nb;
    // p.q partials: as many as the finish reports (per block of its launch),
    // else k_cg_pq's per CG block
    double *pq = c->pq_buf;
    // preconditioner z = pol^T r
    hipLaunchKernelGGL(k_cg_init, dim3(nb), dim3(CG_BLOCK), 0, c->st, n, np, c->pidx, c->type, c->tab_d, c->ef,
                       c->mu_ind, c->dc_R, c->cg_z, c->cg_p, c->dc_d, rz[0], act, wp);
    int cur = 0, launched = 1;
    const int limit = std::max(maxit, minit) + 1;
    auto step = [&]() {
        int keep = c->mf_lo;
        if (lo) c->mf_lo = c->mf_inner;
        c->wpart_ready = wp != nullptr;
        const int npq = induced_field(c, lo, c->cg_p, nullptr, c->dc_g, act, pq);
        c->mf_lo = keep;
        const int nbq = npq > 0 ? npq : nb;
        if (npq == 0)
            hipLaunchKernelGGL(k_cg_pq, dim3(nb), dim3(CG_BLOCK), 0, c->st, n, np, c->pidx, c->type, c->tab_d,
                               c->cg_p, c->dc_g, pq, act);
        hipLaunchKernelGGL(k_cga_update, dim3(nb), dim3(CG_BLOCK), 0, c->st, n, np, nb, nbq, c->pidx, c->type,
                           c->tab_d, c->cg_p, c->dc_g, c->mu_ind, c->ef, c->dc_R, c->cg_z, rz[cur], pq, rz[cur ^ 1],
                           rm, act);
        hipLaunchKernelGGL(k_cga_next, dim3(nb), dim3(CG_BLOCK), 0, c->st, n, np, nb, c->pidx, c->cg_z, c->cg_p,
                           rz[cur], rz[cur ^ 1], rm, c->ctl, c->scal, maxit, minit, c->p.conv_tol_ind, wp);
        cur ^= 1;
        launched++;
    };
    const int hint = refine ? c->hint_ref : c->hint_cg;
    int first = std::min(limit, 1 + std::max(1, hint > 0 ? hint : c->cg_first));
    while (launched < first) step();
    for (;;) {
        int rc = sync_scal(c);
        if (rc) return rc;
        if (!c->ctl_h[C_SCF_ACTIVE] || launched >= limit) break;
        step();
    }
    // u_pol of the final dipoles (pim.c:253,299), its total and the history
    // store in one launch
    hipLaunchKernelGGL(k_upol_hist, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->type, c->tab_d, c->mu_ind,
                       c->upol_i, c->hist, c->chk_part, c->chk_cnt, c->scal);
    scf_finish(c, false);
    if (anchor) c->n_scf_inner += 1;
    else c->n_scf_fp64 += 1;
    if (lo) c->n_scf_inner += std::max(0, c->last_nscf - 1);
    else c->n_scf_fp64 += std::max(0, c->last_nscf - 1);
    if (refine) c->hint_ref = std::max(1, c->last_nscf - 1);
    else c->hint_cg = std::max(1, c->last_nscf - 1);
    if (c->last_rmsd > c->p.conv_tol_ind) return MDFF_ERR_SCF;
    return MDFF_OK;
}

// FP64 refinement of the mixed anchored CG's dipoles: the static field in
// FP64 (FP64 tables of all ions; the FP64 pair tensors were built with the
// FP32 ones), the FP64 field of the converged dipoles as a new anchor, and
// the loop condition of pim.c:236 on that FP64 residual; while it is above
// the tolerance, conjugate-gradient steps in the low precision carry the
// correction (each applies T to the search direction only, whose size is of
// the order of the residual, so their rounding stays far below it).  The
// reported iterations are those of both passes; the residual is the FP64-
// anchored one.
static int scf_refine_fp64(mdff_ctx *c, int istep)
{
    const int first = c->last_nscf;
    launch_eikr(c, 0, 0, nullptr, true);
    static_field(c, 0);
    int rc = scf_loop_anchored(c, istep, true);
    c->last_nscf += first;
    return rc;
}

// the mixed form's SCF: anchored CG (symmetric polarisabilities), else the
// reference's Jacobi schedule with FP64 fields (scf_loop)
static int scf_mixed(mdff_ctx *c, int istep)
{
    if (!scf_anchored_ok(c)) return scf_loop(c, istep);
    int rc = scf_loop_anchored(c, istep);
    const bool run = c->p.max_scf_pim_iter > 0 || c->p.min_scf_pim_iter > 0;
    if (c->scf_refine && run && (rc == MDFF_OK || rc == MDFF_ERR_SCF)) rc = scf_refine_fp64(c, istep);
    return rc;
}

// scfkO, momentpolaSCFkO pim.c:457-603 (FP64 in both forms): the SCF loop
// of momentpolaSCF whose update is mu += dmu from the inner loop of
// induced_moment_inner_kO (pim.c:647-748) on the residual field
// f = E - mu diag(pol^-1), f = E_stat to start; the ASPC predictor only
// when istep > 1, else the first update starts from zero dipoles.  The
// inner sweeps run gated on the device (at most 21 per update); the host
// polls the outer loop condition as scf_loop does.
static void ko_update(mdff_ctx *c, const int *act)
{
    int n = c->n, np = c->npol;
    const int nb = cdiv(np, CG_BLOCK);
    double *in_part = c->cg_part, *ext_part = c->cg_part + nb;
    const int *kact = c->ctl + C_KO_ACTIVE;
    const double om = c->p.omegakO;
    const double a2 = KO_ALPHA * KO_ALPHA;
    const double sself = 2.0 * (2.0 * (KO_ALPHA / sqrt(MDFF_PI)) * a2 / 3.0);   // 2 selfa2 at alpha 0.001
    hipLaunchKernelGGL(k_ko_start, dim3(nb), dim3(CG_BLOCK), 0, c->st, n, np, c->pidx, c->type, c->tab_d, c->ko_f,
                       c->ko_dmu, om, ext_part, c->ctl);
    unsigned blocks = cdiv((size_t)n * 16, 256);
    for (int k = 0; k < 21; k++) {
        hipLaunchKernelGGL((k_t2_field<double, 16>), dim3(blocks), dim3(256), 0, c->st, n, c->capP, np, c->pinv,
                           c->cntp, c->idxp, c->t2ko, c->ko_dmu, c->dc_g, kact);
        hipLaunchKernelGGL(k_ko_sweep, dim3(nb), dim3(CG_BLOCK), 0, c->st, n, np, c->pidx, c->type, c->tab_d,
                           c->ko_f, c->dc_g, sself, om, c->ko_dmu, in_part, c->ctl);
        hipLaunchKernelGGL(k_ko_check, dim3(1), dim3(CG_BLOCK), 0, c->st, nb, np, in_part, ext_part, c->ctl);
    }
    hipLaunchKernelGGL(k_cg_apply, dim3(nb), dim3(CG_BLOCK), 0, c->st, n, np, c->pidx, c->ko_dmu, c->mu_ind, act);
}

static int scf_loop_ko(mdff_ctx *c, int istep)
{
    int n = c->n;
    const int *act = c->ctl + C_SCF_ACTIVE;
    int maxit = c->p.max_scf_pim_iter, minit = c->p.min_scf_pim_iter;
    if (!c->t2_ok) return MDFF_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(k_scf_begin, dim3(1), dim3(1), 0, c->st, c->ctl, (maxit > 0 || minit > 0) ? 1 : 0);
    set_int(c, C_KO_TOTAL, 0);
    AspcArgs asp;
    if (aspc_args(c, istep, &asp)) return MDFF_ERR_CONFIG;
    const bool predict = istep > 1;
    // get_dipoles hands over zero dipoles (coulombic.c:90-96); the residual
    // field of the first update is the static field (pim.c:508-514)
    if (!predict) hipLaunchKernelGGL(k_zero, dim3(cdiv(3 * n, 256)), dim3(256), 0, c->st, c->mu_ind, (size_t)3 * n);
    HIPCHK(hipMemcpyAsync(c->ko_f, c->ef_stat, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->ef, c->ef_stat, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, c->st));
    int launched = 0;
    const int limit = std::max(maxit, minit) + 1;
    while (launched < limit) {
        int chunk = launched == 0 ? std::max(std::max(1, std::min(minit, c->scf_chunk)), c->hint_ko) : c->scf_chunk;
        for (int it = 0; it < chunk && launched < limit; it++, launched++) {
            const int first = launched == 0;
            if (first && predict) {
                hipLaunchKernelGGL(k_scf_mu, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, 1, asp, c->type, c->tab_d,
                                   c->hist, c->ef, c->mu_ind, c->upol_i, act);
            } else {
                ko_update(c, act);
                hipLaunchKernelGGL(k_upol, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->type, c->tab_d,
                                   c->mu_ind, c->upol_i);
            }
            induced_field(c, 0, c->mu_ind, c->ef_stat, c->ef, act);
            if (first)
                hipLaunchKernelGGL(k_scf_correct, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, asp, c->type,
                                   c->tab_d, c->ef, c->mu_ind, act);
            // rmsd, u_pol of this iteration's dipoles and the loop condition;
            // then the residual field of the next update
            scf_check(c, c->upol_i, maxit, minit);
            hipLaunchKernelGGL(k_ko_resid, dim3(cdiv(c->npol, 256)), dim3(256), 0, c->st, n, c->npol, c->pidx,
                               c->type, c->tab_d, c->ef, c->mu_ind, c->ko_f, act);
        }
        int rc = sync_scal(c);
        if (rc) return rc;
        if (!c->ctl_h[C_SCF_ACTIVE]) break;
    }
    scf_finish(c);
    c->n_scf_fp64 += c->last_nscf;
    c->n_scf_inner += c->ctl_h[C_KO_TOTAL];
    c->hint_ko = c->last_nscf;
    if (c->last_rmsd > c->p.conv_tol_ind) return MDFF_ERR_SCF;
    return MDFF_OK;
}

// final multipole_ES of engforce (field.c:381-383): forces and stress
static void coulomb_force(mdff_ctx *c, int lo)
{
    int n = c->n;
    int ld = c->ldip_sticky || c->ldip_static;
    // real space on the second stream (after mu_tot), k space here; joined
    // before the forces are assembled
    aux_fork(c);
    {
        OnAux g(c);
        DirArgs a = dir_args(c, c->q, c->mu_tot, c->ldmp, 0, nullptr);
        launch_dir(c, a, lo, 0, 0, 0, 1);
        colsum(c, c->uw_dir, n, n, 7, c->scal + S_U_DIR);
    }
    int C = ld ? 4 : 1;
    int NCH = ld ? 9 : 3;
    const int done = c->tot_done;
    c->tot_done = 0;
    if (done == 2 && C == 4) {
        c->wpart_ready = 1;
        c->wpart_nb = cdiv(n, 256);
    }
    rec_part(c, 0, lo, C, NCH, c->q, c->mu_tot, c->q, c->mu_tot, 0, 1, 1, nullptr, nullptr);
    if (!done)
        hipLaunchKernelGGL(k_self_terms, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->q, c->mu_tot, c->selfbuf);
    // the reciprocal-space energy/stress terms and the self terms, one launch
    hipLaunchKernelGGL(k_colsum2, dim3(7 + 2), dim3(1024), 0, c->st, c->ek, (int)c->nkf, (int)c->nkf, 7,
                       c->scal + S_U_REC, c->selfbuf, n, n, c->scal + S_QSQ);
}

// FP64 residual of the mixed SCF's final dipoles: the static field and the
// induced field formed entirely in FP64 (FP64 tables, pair tensors and
// matrix-core sums), then sum over the polarisable ions of
// (mu/pol - E)^2 as get_rmsd_mu (pim.c:431-449) into S_RMSD_CHK.  Runs
// after the SCF, before the force call; the mixed ef_stat is no longer
// needed then.
static int scf_fp64_residual(mdff_ctx *c)
{
    int n = c->n;
    launch_eikr(c, 0, 0, nullptr, true);
    int rc = build_t2(c, true);
    if (rc) return rc;
    static_field(c, 0);
    induced_field(c, 0, c->mu_ind, c->ef_stat, c->dc_g, nullptr);
    hipLaunchKernelGGL(k_resid2, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->type, c->tab_d, c->mu_ind, c->dc_g,
                       c->ke);
    colsum(c, c->ke, n, n, 1, c->scal + S_RMSD_CHK);
    return MDFF_OK;
}

static int engforce_impl(mdff_ctx *c, int istep)
{
    int n = c->n;
    int rc = MDFF_OK;
    if (c->have_events) HIPCHK(hipEventRecord(c->ev[0], c->st));
    hipLaunchKernelGGL(k_frac, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->cd, c->r, c->r + n, c->r + 2 * n, c->s);
    // the short-range pair forces on the second stream, beside the k-space
    // tables and sums; joined before the forces are assembled
    aux_fork(c);
    {
        OnAux g(c);
        hipLaunchKernelGGL(k_zero, dim3(cdiv(3 * n, 256)), dim3(256), 0, c->st, c->f_sr, (size_t)3 * n);
        short_range(c);
    }
    if (c->have_events) HIPCHK(hipEventRecord(c->ev[1], c->st));
    c->last_nscf = 0;
    if (c->p.lcoulombic) {
        // k-space tables at the current positions
        int lo = lo_of(c);
        // all ions, and in the same pass the polarisable ions' tables of
        // this precision; the FP64 ones too when the mixed SCF uses them
        const bool pol = c->lpol && c->npol > 0;
        launch_eikr(c, 0, lo, nullptr, pol);
        if (pol && lo && scf_uses_fp64(c)) launch_eikr(c, 1, 0, nullptr);
        // get_dipoles coulombic.c:67-130
        if (c->lpol) {
            if ((rc = build_t2(c, lo && c->scf_refine)) != MDFF_OK) return rc;
            static_field(c, lo);
            if (c->have_events) HIPCHK(hipEventRecord(c->ev[2], c->st));
            rc = c->p.algo_pim == MDFF_PIM_SCFKO ? scf_loop_ko(c, istep)
                 : lo                                 ? scf_mixed(c, istep)
                 : (c->fp64_cg && c->cg_ok)           ? scf_loop_anchored(c, istep)
                                                      : scf_loop(c, istep);
            if (rc != MDFF_OK && rc != MDFF_ERR_SCF) return rc;
            c->chk_done = 0;
            if (lo && c->scf_check > 0 && c->n_force % c->scf_check == 0) {
                int r2 = scf_fp64_residual(c);
                if (r2) return r2;
                c->chk_done = 1;
            }
            c->ldip_sticky = 1;
            // mu_tot with the force call's self terms and, for a split-FP16
            // force sum, its weight maxima
            const bool wt = lo && c->mf_lo == KM_F16X3;
            hipLaunchKernelGGL(k_mu_tot, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->q, c->mu_static, c->mu_ind,
                               c->mu_tot, c->selfbuf, wt ? c->gwp : nullptr);
            c->tot_done = wt ? 2 : 1;
        } else {
            if (c->have_events) HIPCHK(hipEventRecord(c->ev[2], c->st));
            HIPCHK(hipMemcpyAsync(c->mu_tot, c->mu_static, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, c->st));
            hipLaunchKernelGGL(k_zero, dim3(1), dim3(1), 0, c->st, c->scal + S_UPOL, (size_t)1);
        }
        if (c->have_events) HIPCHK(hipEventRecord(c->ev[3], c->st));
        coulomb_force(c, lo);
    } else {
        if (c->have_events) {
            HIPCHK(hipEventRecord(c->ev[2], c->st));
            HIPCHK(hipEventRecord(c->ev[3], c->st));
        }
    }
    aux_join(c);
    hipLaunchKernelGGL(k_assemble_force, dim3(cdiv(3 * n, 256)), dim3(256), 0, c->st, 3 * n, c->f_sr, c->f_dir,
                       c->f_rec, c->p.lcoulombic ? c->u.coul_unit : 0.0, c->f, c->ctl);
    c->n_force++;
    if (c->have_events) HIPCHK(hipEventRecord(c->ev[4], c->st));
    return rc;
}

// thermo from the device scalars (thermo.c:12-33)
static void thermo_from_scal(mdff_ctx *c, mdff_thermo *th)
{
    const double *S = c->scal_h;
    mdff_thermo t;
    memset(&t, 0, sizeof(t));
    double iv = c->cell.inveOmegaPU;
    auto sym = [&](const double *w, double scale, double out[3][3]) {
        out[0][0] = w[0] * scale; out[0][1] = w[1] * scale; out[0][2] = w[2] * scale;
        out[1][0] = w[1] * scale; out[1][1] = w[3] * scale; out[1][2] = w[4] * scale;
        out[2][0] = w[2] * scale; out[2][1] = w[4] * scale; out[2][2] = w[5] * scale;
    };
    if (c->p.lbhmft || c->p.lbhmftd) {
        t.u_bhmftd = S[S_U_BH];
        sym(S + S_W_BH, iv, t.tau_bhmftd);
        t.pvir_bhmftd = (t.tau_bhmftd[0][0] + t.tau_bhmftd[1][1] + t.tau_bhmftd[2][2]) * (1.0 / 3.0);
    }
    if (c->p.lnmlj) {
        t.u_lj = S[S_U_LJ];
        sym(S + S_W_LJ, iv, t.tau_lj);
        t.pvir_lj = (t.tau_lj[0][0] + t.tau_lj[1][1] + t.tau_lj[2][2]) * (1.0 / 3.0);
    }
    if (c->p.lcoulombic) {
        double u_self = -c->selfa * S[S_QSQ] - c->selfa2 * S[S_MUSQ];
        t.u_coul = (S[S_U_DIR] + S[S_U_REC] * c->ttpiV + u_self) * c->u.coul_unit;
        double td[3][3], tr[3][3];
        sym(S + S_W_DIR, 0.5 * iv, td);
        sym(S + S_W_REC, c->ttpiV * iv, tr);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) t.tau_coul[i][j] = td[i][j] * c->u.coul_unit + tr[i][j] * c->u.coul_unit;
        t.pvir_coul = (t.tau_coul[0][0] + t.tau_coul[1][1] + t.tau_coul[2][2]) * (1.0 / 3.0);
        t.u_pol = c->lpol ? S[S_UPOL] * c->u.coul_unit : 0.0;
    }
    t.e_kin = S[S_KIN];
    t.e_nvt = S[S_ENVT];
    double l = 3.0 * (double)c->n;
    t.temp_r = (2.0 * t.e_kin) / (l * c->u.boltz_unit);
    t.u_tot = t.u_lj + t.u_bhmftd + t.u_coul + t.u_pol;
    t.e_tot = t.u_tot + t.e_kin;
    t.h_tot = t.e_tot + t.e_nvt;
    t.pressure = t.pvir_lj + t.pvir_bhmftd + t.pvir_coul + t.temp_r * c->u.boltz_unit * iv;
    t.n_scf = c->last_nscf;
    t.rmsd = c->last_rmsd;
    t.rmsd_fp64 = c->chk_done ? sqrt(S[S_RMSD_CHK] / (double)(c->npol > 0 ? c->npol : 1)) : -1.0;
    t.verlet_updates = c->ctl_h[C_NUPD];
    *th = t;
}

static void kin_reduce(mdff_ctx *c)
{
    colsum(c, c->ke, c->n, c->n, 1, c->scal + S_KIN);
}

// ---------------------------------------------------------------- MD steps
// prop_velocity_verlet integration.c:22-63
static int step_vv(mdff_ctx *c, int istep)
{
    int n = c->n;
    hipLaunchKernelGGL(k_vv_first, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->dt, c->r, c->v, c->f, c->fold,
                       c->invmass);
    int rc = engforce_impl(c, istep);
    if (rc && rc != MDFF_ERR_SCF) return rc;
    hipLaunchKernelGGL(k_vv_second, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->dt, c->v, c->f, c->fold,
                       c->invmass, c->massd, c->ke);
    kin_reduce(c);
    return rc;
}

// nhcn integration.c:140-182 with chain_nhcn :189-294
static void nhc_chain(mdff_ctx *c, bool reduce = true, bool envt = false)
{
    NhcArgs a;
    a.nh = c->p.nhc_n;
    a.mults = c->p.nhc_mults;
    a.yo = c->p.nhc_yosh_order;
    a.dt = c->dt;
    a.kT = c->kT;
    a.L = 3.0 * (double)c->n;
    for (int i = 0; i < 16; i++) {
        a.Q[i] = c->nhcQ[i];
        a.iQ[i] = c->nhcQ[i] != 0.0 ? 1.0 / c->nhcQ[i] : 0.0;
    }
    for (int i = 0; i < 7; i++) a.w[i] = c->yosh[i];
    if (a.nh <= 4)
        hipLaunchKernelGGL(k_nhc_chain<4>, dim3(1), dim3(1), 0, c->st, a, c->scal, envt ? 1 : 0);
    else if (a.nh <= 8)
        hipLaunchKernelGGL(k_nhc_chain<8>, dim3(1), dim3(1), 0, c->st, a, c->scal, envt ? 1 : 0);
    else
        hipLaunchKernelGGL(k_nhc_chain<16>, dim3(1), dim3(1), 0, c->st, a, c->scal, envt ? 1 : 0);
    // with reduce, the scaling kernel also sums the kinetic energy into S_KIN
    // (the SCF-check partials and count are free here)
    if (reduce)
        hipLaunchKernelGGL(k_scale_v, dim3(cdiv(c->n, 256)), dim3(256), 0, c->st, c->n, c->scal + S_NHC_S, c->v,
                           c->massd, c->ke, c->chk_part, c->chk_cnt, c->scal + S_KIN);
    else
        hipLaunchKernelGGL(k_scale_v, dim3(cdiv(c->n, 256)), dim3(256), 0, c->st, c->n, c->scal + S_NHC_S, c->v,
                           c->massd, c->ke, (double *)nullptr, (unsigned *)nullptr, (double *)nullptr);
}

static int step_nhc(mdff_ctx *c, int istep)
{
    int n = c->n;
    // the kinetic energy of the velocities the last step left (formed at its
    // end, unless anything changed the velocities since)
    if (!c->kin_fresh) {
        hipLaunchKernelGGL(k_ke, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->v, c->massd, c->ke);
        kin_reduce(c);
    }
    c->kin_fresh = 0;
    // first half: S_KIN after the scaling is not read before k_kick forms it again
    nhc_chain(c, false);
    hipLaunchKernelGGL(k_kick_drift, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->dt, c->r, c->v, c->f,
                       c->invmass);
    int rc = engforce_impl(c, istep);
    if (rc && rc != MDFF_ERR_SCF) return rc;
    // the kick also sums the kinetic energy into S_KIN
    hipLaunchKernelGGL(k_kick, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->dt, c->v, c->f, c->invmass, c->massd,
                       c->ke, c->chk_part, c->chk_cnt, c->scal + S_KIN);
    // second half, with e_nvt of the advanced chain
    nhc_chain(c, true, true);
    c->kin_fresh = 1;
    return rc;
}

// prop_leap_frog integration.c:93-137
static int step_lf(mdff_ctx *c, int istep)
{
    int n = c->n;
    int rc = engforce_impl(c, istep);
    if (rc && rc != MDFF_ERR_SCF) return rc;
    hipLaunchKernelGGL(k_leapfrog, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->dt, c->r, c->rs, c->v, c->f,
                       c->invmass, c->massd, c->ke);
    kin_reduce(c);
    return rc;
}

// rescale_velocities kinetic.c:83-136 on the device
static void rescale_dev(mdff_ctx *c)
{
    int n = c->n;
    hipLaunchKernelGGL(k_ke, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->v, c->massd, c->ke);
    kin_reduce(c);
    double ratio = (c->integ != MDFF_NVT_NHCN) ? c->dt / c->tauT : -1.0;
    hipLaunchKernelGGL(k_rescale_lambda, dim3(1), dim3(1), 0, c->st, c->scal, n, c->kT, c->u.boltz_unit, ratio);
    hipLaunchKernelGGL(k_scale_v, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->scal + S_LAMBDA, c->v, c->massd,
                       c->ke);
    colsum(c, c->v, n, n, 3, c->scal + S_VMEAN);
    hipLaunchKernelGGL(k_sub_mean, dim3(cdiv(3 * n, 256)), dim3(256), 0, c->st, n, c->scal + S_VMEAN, c->v);
}

// prop_agate integration.c:298-333: the integrator's step with its force call
static int prop_step(mdff_ctx *c, int istep)
{
    switch (c->integ) {
    case MDFF_NVE_LF: return step_lf(c, istep);
    case MDFF_NVT_NHCN: return step_nhc(c, istep);
    default: return step_vv(c, istep);
    }
}

// a frame of the device trajectory: copies on the stream, no host wait
static int traj_append(mdff_ctx *c, int istep)
{
    if (c->traj_n >= c->traj_cap) {
        c->traj_drop++;
        return MDFF_OK;
    }
    const size_t n3 = (size_t)3 * c->n;
    double *fr = c->traj + (size_t)c->traj_n * 3 * n3;
    HIPCHK(hipMemcpyAsync(fr, c->r, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(fr + n3, c->v, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(fr + 2 * n3, c->f, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->st));
    c->traj_step[c->traj_n++] = istep;
    return MDFF_OK;
}

static int md_step_impl(mdff_ctx *c, int istep)
{
    int rc = prop_step(c, istep);
    if (rc && rc != MDFF_ERR_SCF) return rc;
    if (c->p.lverletL) {
        int r2 = build_list(c, 0);
        if (r2) return r2;
    }
    if (c->rescale_allowed && istep < c->p.nequil && c->p.nequilT > 0 && istep % c->p.nequilT == 0) {
        rescale_dev(c);
        c->kin_fresh = 0;
    }
    // the state run_md writes to TRAJFF after this step (md.c:275-277)
    if (c->traj_every > 0 && (istep % c->traj_every == 0 || istep == c->traj_npas) && istep >= c->traj_start) {
        int r3 = traj_append(c, istep);
        if (r3) return r3;
    }
    return rc;
}

// ================================================================= C ABI
extern "C" {

const char *mdff_version(void) { return "mdff20_amd 0.1 (gfx950)"; }

int mdff_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void mdff_tt_damping(double b, double c, double r, double *f, double *fd, int order)
{
    double E[MDFF_MAX_TT_ORDER + 1];
    E[0] = 1.0;
    for (int k = 1; k <= MDFF_MAX_TT_ORDER; k++) E[k] = E[k - 1] / (double)k;
    if (order < 0) order = 0;
    if (order > MDFF_MAX_TT_ORDER) order = MDFF_MAX_TT_ORDER;
    double br = b * r, ex = exp(-br) * c, ff = E[order];
    for (int k = order - 1; k > -1; k--) ff = ff * br + E[k];
    *f = 1.0 - ff * ex;
    double pw = 1.0;
    for (int k = 0; k < order; k++) pw *= br;
    *fd = E[order] * pw * ex * b;
}

void mdff_destroy(mdff_ctx *c)
{
    if (!c) return;
    if (c->st) (void)hipStreamSynchronize(c->st);
    double **dp[] = {&c->r, &c->v, &c->f, &c->fold, &c->rs, &c->s, &c->s0, &c->invmass, &c->massd, &c->q, &c->qz,
                     &c->mu_static, &c->mu_ind, &c->mu_tot, &c->hist, &c->ef_stat, &c->ef, &c->ef_dir, &c->ef_rec,
                     &c->efg, &c->efg_dir, &c->efg_rec, &c->f_sr, &c->f_dir, &c->f_rec, &c->acc_sr, &c->uw_dir,
                     &c->selfbuf, &c->ke, &c->upol_i, &c->rmsd_i, &c->tmp_q, &c->tmp_mu, &c->tmp_qraw, &c->tmp_mur, &c->dc_R, &c->dc_d, &c->dc_g, &c->cg_z, &c->cg_p, &c->cg_part, &c->chk_part, &c->pq_buf, &c->t2dir, &c->vc_part, &c->ko_f, &c->ko_dmu, &c->t2ko, &c->scal,
                     &c->kvec, &c->Ak, &c->kcoe, &c->W, &c->ek, &c->kpart, &c->bpart, &c->ksum, &c->gsc, &c->gwp, &c->gwq};
    for (auto pp : dp) dfree(*pp);
    free_cells(c);
    dfree(c->traj);
    dfree(c->type);
    dfree(c->vc_cnt);
    dfree(c->chk_cnt);
    c->ctl = nullptr;   // inside the scal allocation
    dfree(c->tab_d);
    dfree(c->pidx);
    dfree(c->pinv);
    if (c->gcm) (void)hipFree(c->gcm);
    dfree(c->cntp);
    dfree(c->idxp);
    dfree(c->t2_ovf);
    dfree(c->t2);
    if (c->t2f) (void)hipFree(c->t2f);
    dfree(c->nbr);
    dfree(c->cnt);
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int d = 0; d < 3; d++)
                if (c->tab3[a][b][d]) (void)hipFree(c->tab3[a][b][d]);
    for (int m = 0; m < 4; m++)
        for (int h = 0; h < 2; h++) {
            for (int pl = 0; pl < 2; pl++) {
                if (c->yz[m][pl][h]) (void)hipFree(c->yz[m][pl][h]);
                if (c->yzt[m][pl][h]) (void)hipFree(c->yzt[m][pl][h]);
            }
            if (c->aw[m][h]) (void)hipFree(c->aw[m][h]);
            if (c->wm[m][h]) (void)hipFree(c->wm[m][h]);
        }
    if (c->scal_h) (void)hipHostFree(c->scal_h);   // ctl_h lies inside it
    c->scal_h = nullptr;
    c->ctl_h = nullptr;
    if (c->have_events)
        for (int i = 0; i < 8; i++) (void)hipEventDestroy(c->ev[i]);
    if (c->st2) (void)hipStreamSynchronize(c->st2);
    for (auto &e : c->xev)
        if (e) (void)hipEventDestroy(e);
    if (c->st2) (void)hipStreamDestroy(c->st2);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

int mdff_create(mdff_ctx **out, const mdff_params *p, int nion, const int *typia, const double *A9,
                const double *rx, const double *ry, const double *rz, int precision, int device)
{
    if (!out || !p || nion < 1 || !typia || !A9 || !rx || !ry || !rz) return MDFF_ERR_ARG;
    if (p->ntype < 1 || p->ntype > NT) return MDFF_ERR_ARG;
    if (precision != MDFF_FP64 && precision != MDFF_MIXED) return MDFF_ERR_ARG;
    if (p->lcoulombic && p->algo_pim != MDFF_PIM_SCF && p->algo_pim != MDFF_PIM_SCFKO) {
        int anypol = 0;
        for (int t = 0; t < p->ntype; t++) anypol |= p->lpolar[t];
        if (anypol) return MDFF_ERR_UNSUPPORTED;
    }
    if (p->integrator == MDFF_NVT_NHCN &&
        (p->nhc_n < 2 || p->nhc_n > 16 || p->nhc_mults < 1 ||
         !(p->nhc_yosh_order == 1 || p->nhc_yosh_order == 3 || p->nhc_yosh_order == 5 || p->nhc_yosh_order == 7 ||
           p->nhc_yosh_order == 9)))
        return MDFF_ERR_CONFIG;
    if (p->extrapolate_order < 0 || p->extrapolate_order > MDFF_MAX_EXTRAPOLATE_ORDER) return MDFF_ERR_CONFIG;
    for (int i = 0; i < nion; i++)
        if (typia[i] < 0 || typia[i] >= p->ntype) return MDFF_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MDFF_ERR_DEVICE;
    if (device < 0 || device >= ndev) return MDFF_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return MDFF_ERR_DEVICE;

    mdff_ctx *c = new mdff_ctx();
    c->p = *p;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            c->ncu = prop.multiProcessorCount;
    }
    c->n = nion;
    c->ntype = p->ntype;
    c->precision = precision;
    c->device = device;
    mdff_units_get(p, &c->u);
    double A[3][3];
    for (int i = 0; i < 9; i++) A[i / 3][i % 3] = A9[i];
    mdff_lattice(A, c->u.press_unit, &c->cell);
    memcpy(c->cd.A, c->cell.A, sizeof(c->cd.A));
    memcpy(c->cd.B, c->cell.B, sizeof(c->cd.B));
    c->typ.assign(typia, typia + nion);
    // flags (field.c:195-219)
    for (int t = 0; t < c->ntype; t++) {
        if (p->qit[t] != 0.0) c->lqch = 1;
        for (int j = 0; j < 3; j++) {
            if (p->dipit[t][j] != 0.0) c->ldip_static = 1;
            for (int k = 0; k < 3; k++)
                if (p->polit[t][j][k] != 0.0) c->lpol = 1;
        }
        if (p->lpolar[t]) c->lpim = 1;
    }
    if (!p->lcoulombic) c->lpol = 0;
    for (int t = 0; t < c->ntype; t++)
        for (int i = 0; i < c->ntype; i++)
            for (int k = 0; k < c->ntype; k++)
                if (p->lpoldamping[t][i][k]) c->ldmp = 1;
    c->lnonbonded = p->lnmlj || p->lbhmft || p->lbhmftd;
    c->srcutsq = p->cutshortrange * p->cutshortrange;
    c->lrcutsq = p->cutlongrange * p->cutlongrange;
    c->alpha = p->alphaES;
    memcpy(c->kES, p->kES, sizeof(c->kES));
    if (p->lcoulombic && p->lautoES) mdff_auto_ewald(p, &c->cell, &c->alpha, c->kES);
    double a2 = c->alpha * c->alpha;
    c->selfa = c->alpha / sqrt(MDFF_PI);
    c->selfa2 = 2.0 * c->selfa * a2 / 3.0;
    c->ttpiV = 2.0 * MDFF_TPI * c->cell.inveOmega;
    c->tfpiV = 2.0 * c->ttpiV;
    // md (check_md md.c:120-141)
    c->integ = (p->integrator == MDFF_NVE_LF && p->nequil > 0) ? MDFF_NVE_VV : p->integrator;
    c->rescale_allowed = (c->integ == MDFF_NVE_VV);
    double tauT = p->tauTberendsen;
    if (p->nequil > 0 && tauT == 0.0) tauT = p->dt;
    c->kT = p->temp * c->u.boltz_unit;
    c->dt = p->dt * c->u.time_unit;
    c->tauT = tauT * c->u.time_unit;
    c->timesca = p->timesca_thermo * c->u.time_unit;
    for (int i = 0; i < 16; i++) c->nhcQ[i] = c->timesca * c->timesca * c->kT;
    c->nhcQ[0] *= 3.0 * (double)nion;
    memset(c->yosh, 0, sizeof(c->yosh));
    switch (p->nhc_yosh_order) {
    case 1: c->yosh[0] = 1.0; break;
    case 3:
        c->yosh[0] = 1.0 / (2.0 - pow(2.0, 1.0 / 3.0));
        c->yosh[1] = 1.0 - 2.0 * c->yosh[0];
        c->yosh[2] = c->yosh[0];
        break;
    case 5:
        c->yosh[0] = 1.0 / (4.0 - pow(4, 1.0 / 3.0));
        c->yosh[1] = c->yosh[2] = c->yosh[3] = c->yosh[0];
        c->yosh[4] = 1.0 - 4.0 * c->yosh[0];
        break;
    case 7:
        c->yosh[0] = .78451361047756; c->yosh[1] = .235573213359357; c->yosh[2] = -1.17767998417887;
        c->yosh[3] = 1.0 - 2.0 * (c->yosh[0] + c->yosh[1] + c->yosh[2]);
        c->yosh[4] = c->yosh[2]; c->yosh[5] = c->yosh[1]; c->yosh[6] = c->yosh[0];
        break;
    default: break;
    }
    build_tables(c);

    int rc = MDFF_OK;
    int n = nion;
#define ALLOC(ptr, cnt) do { if ((rc = dalloc(&(ptr), (cnt))) != MDFF_OK) { mdff_destroy(c); return rc; } } while (0)
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) { delete c; return MDFF_ERR_DEVICE; }
    if (hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->xev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->xev[1], hipEventDisableTiming) != hipSuccess) {
        mdff_destroy(c);
        return MDFF_ERR_DEVICE;
    }
    ALLOC(c->r, 3 * n); ALLOC(c->v, 3 * n); ALLOC(c->f, 3 * n); ALLOC(c->fold, 3 * n); ALLOC(c->rs, 3 * n);
    ALLOC(c->s, 3 * n); ALLOC(c->s0, 3 * n); ALLOC(c->type, n); ALLOC(c->invmass, n); ALLOC(c->massd, n);
    ALLOC(c->q, n); ALLOC(c->qz, n); ALLOC(c->mu_static, 3 * n); ALLOC(c->mu_ind, 3 * n); ALLOC(c->mu_tot, 3 * n);
    ALLOC(c->hist, 18 * (size_t)n); ALLOC(c->ef_stat, 3 * n); ALLOC(c->ef, 3 * n); ALLOC(c->ef_dir, 3 * n);
    ALLOC(c->ef_rec, 3 * n); ALLOC(c->efg, 9 * n); ALLOC(c->efg_dir, 9 * n); ALLOC(c->efg_rec, 9 * n);
    ALLOC(c->f_sr, 3 * n); ALLOC(c->f_dir, 3 * n); ALLOC(c->f_rec, 3 * n); ALLOC(c->acc_sr, 14 * n);
    ALLOC(c->uw_dir, 7 * n); ALLOC(c->selfbuf, 2 * n); ALLOC(c->ke, n); ALLOC(c->upol_i, n); ALLOC(c->rmsd_i, n);
    ALLOC(c->tmp_q, n); ALLOC(c->tmp_mu, 3 * n); ALLOC(c->tmp_qraw, n); ALLOC(c->tmp_mur, 3 * n);
    ALLOC(c->dc_R, 3 * n); ALLOC(c->dc_d, 3 * n); ALLOC(c->dc_g, 3 * n);
    ALLOC(c->cg_z, 3 * n); ALLOC(c->cg_p, 3 * n); ALLOC(c->cg_part, 4 * (size_t)cdiv(n, CG_BLOCK));
    ALLOC(c->chk_part, 2 * (size_t)cdiv(n, 256));
    ALLOC(c->pq_buf, (size_t)cdiv(n, 4)); ALLOC(c->t2dir, 3 * n);
    ALLOC(c->vc_part, 2 * (size_t)cdiv(n, 256)); ALLOC(c->vc_cnt, 1); ALLOC(c->chk_cnt, 1);
    ALLOC(c->ko_f, 3 * n); ALLOC(c->ko_dmu, 3 * n);
    // scal [S_COUNT doubles] then ctl [C_COUNT ints] in one allocation, so
    // that one copy brings both to the host
    ALLOC(c->scal, S_COUNT + ((size_t)C_COUNT * sizeof(int) + sizeof(double) - 1) / sizeof(double));
    c->ctl = (int *)(c->scal + S_COUNT);
    ALLOC(c->tab_d, 1);
    if (hipHostMalloc((void **)&c->scal_h, sizeof(double) * S_COUNT + sizeof(int) * C_COUNT) != hipSuccess) {
        mdff_destroy(c);
        return MDFF_ERR_NOMEM;
    }
    c->ctl_h = (int *)(c->scal_h + S_COUNT);
    memset(c->ctl_h, 0, sizeof(int) * C_COUNT);
    memset(c->scal_h, 0, sizeof(double) * S_COUNT);
    // per-ion host arrays (init_config config.c:124-198)
    std::vector<double> hm(n), him(n), hq(n), hmu(3 * n, 0.0);
    std::vector<int> hp;
    c->mass.resize(n);
    for (int i = 0; i < n; i++) {
        int t = typia[i];
        hm[i] = p->massit[t];
        him[i] = 1.0 / p->massit[t];
        c->mass[i] = hm[i];
        hq[i] = p->lcoulombic ? p->qit[t] : 0.0;
        if (p->lcoulombic)
            for (int k = 0; k < 3; k++) hmu[(size_t)k * n + i] = p->dipit[t][k];
        if (c->lpol && p->lpolar[t]) hp.push_back(i);
    }
    c->npol = (int)hp.size();
    std::vector<double> hr(3 * n);
    memcpy(hr.data(), rx, sizeof(double) * n);
    memcpy(hr.data() + n, ry, sizeof(double) * n);
    memcpy(hr.data() + 2 * n, rz, sizeof(double) * n);
    if (hipMemcpy(c->r, hr.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->type, typia, sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->massd, hm.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->invmass, him.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->q, hq.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->mu_static, hmu.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->tab_d, &c->tab, sizeof(TypeTab), hipMemcpyHostToDevice) != hipSuccess) {
        mdff_destroy(c);
        return MDFF_ERR_DEVICE;
    }
    if (c->npol > 0) {
        ALLOC(c->pidx, c->npol);
        if (hipMemcpy(c->pidx, hp.data(), sizeof(int) * c->npol, hipMemcpyHostToDevice) != hipSuccess) {
            mdff_destroy(c);
            return MDFF_ERR_DEVICE;
        }
        std::vector<int> hinv(n, -1);
        for (int k = 0; k < c->npol; k++) hinv[hp[k]] = k;
        ALLOC(c->pinv, n);
        ALLOC(c->cntp, c->npol);
        ALLOC(c->t2_ovf, 1);
        if (hipMemcpy(c->pinv, hinv.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess) {
            mdff_destroy(c);
            return MDFF_ERR_DEVICE;
        }
    }
    // k mesh (init_kspace kspace.c:14-41, set_param_kmesh :64-102) on the full
    // (nx, ny, nz) box with nx >= 0; entries outside the half space get Ak = 0
    if (p->lcoulombic) {
        c->Kx = c->kES[0]; c->Ky = c->kES[1]; c->Kz = c->kES[2];
        if (c->Kx < 0 || c->Ky < 0 || c->Kz < 0) { mdff_destroy(c); return MDFF_ERR_CONFIG; }
        c->P = (2 * c->Ky + 1) * (2 * c->Kz + 1);
        c->nkf = (size_t)(c->Kx + 1) * c->P;
        std::vector<double> kv(3 * c->nkf), ak(c->nkf, 0.0), kc(c->nkf, 0.0);
        int nk = 0;
        for (int nx = 0; nx <= c->Kx; nx++)
            for (int ny = -c->Ky; ny <= c->Ky; ny++)
                for (int nz = -c->Kz; nz <= c->Kz; nz++) {
                    size_t idx = (size_t)nx * c->P + (size_t)(ny + c->Ky) * (2 * c->Kz + 1) + (nz + c->Kz);
                    const double(*B)[3] = c->cell.B;
                    double kx = MDFF_TPI * ((double)nx * B[0][0] + (double)ny * B[0][1] + (double)nz * B[0][2]);
                    double ky = MDFF_TPI * ((double)nx * B[1][0] + (double)ny * B[1][1] + (double)nz * B[1][2]);
                    double kz = MDFF_TPI * ((double)nx * B[2][0] + (double)ny * B[2][1] + (double)nz * B[2][2]);
                    kv[3 * idx] = kx; kv[3 * idx + 1] = ky; kv[3 * idx + 2] = kz;
                    bool in_half = (nx > 0) || (nx == 0 && ny > 0) || (nx == 0 && ny == 0 && nz > 0);
                    double kk = kx * kx + ky * ky + kz * kz;
                    if (in_half && kk != 0.0) {
                        ak[idx] = exp(-kk * 0.25 / a2) / kk;
                        kc[idx] = 2.0 * (1.0 / kk + 1.0 / a2 / 4.0);
                        nk++;
                    }
                }
        c->nk = nk;
        ALLOC(c->kvec, 3 * c->nkf); ALLOC(c->Ak, c->nkf); ALLOC(c->kcoe, c->nkf); ALLOC(c->W, 2 * c->nkf);
        ALLOC(c->ek, 7 * c->nkf);
        ALLOC(c->gsc, 16 + 6 * (size_t)(c->Kx + 1));
        ALLOC(c->gwp, 4 * (size_t)cdiv(c->n, 256));
        ALLOC(c->gwq, (size_t)cdiv(c->n, 256));
        if (hipMalloc((void **)&c->gcm, sizeof(unsigned long long) * 6 * (size_t)(c->Kx + 1)) != hipSuccess) {
            mdff_destroy(c);
            return MDFF_ERR_NOMEM;
        }
        if (hipMemcpy(c->kvec, kv.data(), sizeof(double) * 3 * c->nkf, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->Ak, ak.data(), sizeof(double) * c->nkf, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->kcoe, kc.data(), sizeof(double) * c->nkf, hipMemcpyHostToDevice) != hipSuccess) {
            mdff_destroy(c);
            return MDFF_ERR_DEVICE;
        }
        size_t dims[3] = {(size_t)(c->Kx + 1), (size_t)(2 * c->Ky + 1), (size_t)(2 * c->Kz + 1)};
        for (int lo = 0; lo < 2; lo++) {
            if (lo && !lo_of(c)) continue;
            size_t rs = lo ? sizeof(float) : sizeof(double);
            for (int polar = 0; polar < 2; polar++) {
                size_t ns = polar ? (size_t)c->npol : (size_t)n;
                if (ns == 0) continue;
                for (int d = 0; d < 3; d++)
                    if (hipMalloc(&c->tab3[lo][polar][d], 2 * dims[d] * ns * rs) != hipSuccess) {
                        mdff_destroy(c);
                        return MDFF_ERR_NOMEM;
                    }
            }
        }
        // split-K over ions for the forward sum, split over (nx,ny) rows for
        // the back-projection: aim at >= 2048 blocks on 256 CUs
        int ntile = (c->P + 255) / 256;
        int base = (c->Kx + 1) * ntile;
        c->nchunk = std::max(1, std::min(64, (2048 + base - 1) / base));
        c->chunk = std::max(32, ((n + c->nchunk - 1) / c->nchunk + 31) / 32 * 32);
        c->nchunk = (n + c->chunk - 1) / c->chunk;
        if (c->npol > 0) {
            c->nchunk_p = std::max(1, std::min(64, (2048 + base - 1) / base));
            c->chunk_p = std::max(32, ((c->npol + c->nchunk_p - 1) / c->nchunk_p + 31) / 32 * 32);
            c->nchunk_p = (c->npol + c->chunk_p - 1) / c->chunk_p;
        }
        int nrows = (c->Kx + 1) * (2 * c->Ky + 1);
        int nblk = (n + 63) / 64;
        c->nsplit = std::max(1, std::min(nrows, (2048 + nblk - 1) / nblk));
        c->rows_per_split = (nrows + c->nsplit - 1) / c->nsplit;
        c->nsplit = (nrows + c->rows_per_split - 1) / c->rows_per_split;
        int C = 4;
        ALLOC(c->kpart, (size_t)std::max(64, std::max(c->nchunk, c->nchunk_p)) * c->nkf * 2 * C);
        ALLOC(c->ksum, c->nkf * 2 * C);
        c->nsplit_max = std::max(c->nsplit, 16);
        ALLOC(c->bpart, (size_t)c->nsplit_max * 9 * n);
        // padded sizes of the stored-operand sums (gemm_kernels.h)
        c->nselp[0] = (n + 127) / 128 * 128;
        c->nselp[1] = (c->npol + 127) / 128 * 128;
        c->Pp = (c->P + 255) / 256 * 256;
        c->Pk = (c->P + 31) / 32 * 32;   // 2 Pk: whole pairs of K steps
    }
    // neighbour list capacity: all pairs when no list, else an estimate
    // from the number density grown on overflow
    {
        double rcut = std::max(p->cutshortrange, p->cutlongrange) + p->skindiff;
        double est = 4.0 / 3.0 * MDFF_PI * rcut * rcut * rcut * ((double)n * c->cell.inveOmega);
        int capn = p->lverletL ? (int)(est * 1.3) + 64 : n;
        capn = std::min(capn, ((n + 63) / 64) * 64);
        capn = std::max(capn, 64);
        c->cap = (capn + 63) / 64 * 64;
        ALLOC(c->nbr, (size_t)n * c->cap);
        ALLOC(c->cnt, n);
    }
#undef ALLOC
    for (int i = 0; i < 8; i++)
        if (hipEventCreate(&c->ev[i]) != hipSuccess) { mdff_destroy(c); return MDFF_ERR_DEVICE; }
    if (!g_back_attrs(c)) {
        mdff_destroy(c);
        return MDFF_ERR_DEVICE;
    }
    c->have_events = 1;
    // first list build (main calls check_verletlist before MD, mdff20.c:87)
    set_int(c, C_FIRST, 1);
    rc = build_list(c, 1);
    if (rc) { mdff_destroy(c); return rc; }
    // every zero fill of the set-up (null stream) and the first build are done
    if (hipDeviceSynchronize() != hipSuccess) { mdff_destroy(c); return MDFF_ERR_DEVICE; }
    *out = c;
    return MDFF_OK;
}

int mdff_set_kspace_shape(mdff_ctx *c, int cw64, int cw32, int fwd_blocks, int back_blocks)
{
    if (!c) return MDFF_ERR_ARG;
    int cw[2] = {cw64, cw32};
    for (int k = 0; k < 2; k++)
        if (cw[k] && cw[k] != 1 && cw[k] != 2 && cw[k] != 4) return MDFF_ERR_ARG;
    if (fwd_blocks < 0 || back_blocks < 0) return MDFF_ERR_ARG;
    for (int k = 0; k < 2; k++)
        if (cw[k]) c->g_cw[k] = cw[k];
    if (fwd_blocks) c->g_fblocks = fwd_blocks;
    if (back_blocks) c->g_bblocks = back_blocks;
    c->gtune.clear();   // timed block targets belong to the old shapes
    return MDFF_OK;
}

int mdff_set_scf_anchor(mdff_ctx *c, int anchor_mode)
{
    if (!c) return MDFF_ERR_ARG;
    if (!(anchor_mode == 0 || anchor_mode == KM_F32 || anchor_mode == KM_F16X3))
        return MDFF_ERR_ARG;
    c->mf_anchor = anchor_mode;
    return MDFF_OK;
}

int mdff_set_kspace_modes(mdff_ctx *c, int eval_mode, int inner_mode)
{
    if (!c) return MDFF_ERR_ARG;
    auto ok = [](int m) { return m == 0 || m == KM_F32 || m == KM_F16X3; };
    if (!ok(eval_mode) || !ok(inner_mode)) return MDFF_ERR_ARG;
    if (eval_mode) c->mf_lo = eval_mode;
    if (inner_mode) c->mf_inner = inner_mode;
    return MDFF_OK;
}

int mdff_set_tuning(mdff_ctx *c, int lanes, int katoms, int ksplit)
{
    if (!c) return MDFF_ERR_ARG;
    if (lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64) c->lanes = lanes;
    else if (lanes != 0) return MDFF_ERR_ARG;
    // katoms: reciprocal-space kernel family.  -1 vector ALU; 1 the
    // matrix-core sums with split-FP16 SCF steps; 2 the same with FP32 steps;
    // 4 as 1 with the partial sums, reduction and pair-tensor finish in
    // separate kernels.
    if (katoms == 1 || katoms == 4) {
        c->use_mfma = 1; c->mf_inner = KM_F16X3;
        c->no_fuse = katoms == 4;
    }
    else if (katoms == 2) { c->use_mfma = 1; c->mf_inner = KM_F32; c->no_fuse = 0; }
    else if (katoms == -1) c->use_mfma = 0;
    else if (katoms != 0) return MDFF_ERR_ARG;
    if (ksplit > 0) c->mf_blocks = ksplit;
    return MDFF_OK;
}

static int put3(mdff_ctx *c, double *dst, const double *x, const double *y, const double *z)
{
    int n = c->n;
    HIPCHK(hipMemcpyAsync(dst, x, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(dst + n, y, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(dst + 2 * n, z, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return MDFF_OK;
}

static int get3(mdff_ctx *c, const double *src, double *x, double *y, double *z)
{
    int n = c->n;
    if (x) HIPCHK(hipMemcpyAsync(x, src, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    if (y) HIPCHK(hipMemcpyAsync(y, src + n, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    if (z) HIPCHK(hipMemcpyAsync(z, src + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return MDFF_OK;
}

int mdff_set_positions(mdff_ctx *c, const double *rx, const double *ry, const double *rz)
{
    if (!c || !rx || !ry || !rz) return MDFF_ERR_ARG;
    return put3(c, c->r, rx, ry, rz);
}

int mdff_set_velocities(mdff_ctx *c, const double *vx, const double *vy, const double *vz)
{
    if (!c || !vx || !vy || !vz) return MDFF_ERR_ARG;
    return put3(c, c->v, vx, vy, vz);
}

int mdff_get_positions(mdff_ctx *c, double *rx, double *ry, double *rz) { return c ? get3(c, c->r, rx, ry, rz) : MDFF_ERR_ARG; }
int mdff_get_velocities(mdff_ctx *c, double *vx, double *vy, double *vz) { return c ? get3(c, c->v, vx, vy, vz) : MDFF_ERR_ARG; }
int mdff_get_forces(mdff_ctx *c, double *fx, double *fy, double *fz) { return c ? get3(c, c->f, fx, fy, fz) : MDFF_ERR_ARG; }

static int get_aos3(mdff_ctx *c, const double *src, double *out)
{
    int n = c->n;
    std::vector<double> h(3 * n);
    HIPCHK(hipMemcpyAsync(h.data(), src, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) out[3 * i + k] = h[(size_t)k * n + i];
    return MDFF_OK;
}

int mdff_get_dipoles(mdff_ctx *c, double *mu) { return (c && mu) ? get_aos3(c, c->mu_tot, mu) : MDFF_ERR_ARG; }

int mdff_set_trajectory(mdff_ctx *c, int tprint, int tstart, int npas, int capacity)
{
    if (!c || tprint < 0 || capacity < 0 || (tprint > 0 && capacity < 1)) return MDFF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->st));
    dfree(c->traj);
    c->traj_every = 0;
    c->traj_cap = c->traj_n = c->traj_drop = 0;
    c->traj_step.clear();
    if (tprint == 0) return MDFF_OK;
    if (hipMalloc((void **)&c->traj, sizeof(double) * 9 * (size_t)c->n * (size_t)capacity) != hipSuccess) {
        c->traj = nullptr;
        return MDFF_ERR_NOMEM;
    }
    c->traj_step.assign(capacity, 0);
    c->traj_every = tprint;
    c->traj_start = tstart;
    c->traj_npas = npas;
    c->traj_cap = capacity;
    return MDFF_OK;
}

int mdff_traj_info(mdff_ctx *c, int out[3])
{
    if (!c || !out) return MDFF_ERR_ARG;
    out[0] = c->traj_n;
    out[1] = c->traj_cap;
    out[2] = c->traj_drop;
    return MDFF_OK;
}

int mdff_traj_frame(mdff_ctx *c, int k, int *istep, double *r[3], double *v[3], double *f[3])
{
    if (!c || k < 0 || k >= c->traj_n) return MDFF_ERR_ARG;
    const size_t n = (size_t)c->n;
    const double *fr = c->traj + (size_t)k * 9 * n;
    double **dst[3] = {r, v, f};
    for (int q = 0; q < 3; q++)
        if (dst[q])
            for (int a = 0; a < 3; a++)
                if (dst[q][a])
                    HIPCHK(hipMemcpyAsync(dst[q][a], fr + (3 * q + a) * n, sizeof(double) * n, hipMemcpyDeviceToHost,
                                          c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (istep) *istep = c->traj_step[k];
    return MDFF_OK;
}

int mdff_traj_clear(mdff_ctx *c)
{
    if (!c) return MDFF_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->st));
    c->traj_n = c->traj_drop = 0;
    return MDFF_OK;
}

// the state the RDF unit reads: positions, cell, the trajectory's frames
int mdff_engine_view(mdff_ctx *c, EngineView *v)
{
    if (!c || !v) return MDFF_ERR_ARG;
    v->n = c->n;
    v->device = c->device;
    v->st = c->st;
    v->r = c->r;
    v->type = c->type;
    v->ntype = c->ntype;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            v->A[i][j] = c->cd.A[i][j];
            v->B[i][j] = c->cd.B[i][j];
        }
    v->frames = c->traj;
    v->nframes = c->traj_n;
    v->frame_doubles = 9 * (size_t)c->n;
    return MDFF_OK;
}
int mdff_get_field(mdff_ctx *c, double *ef) { return (c && ef) ? get_aos3(c, c->ef, ef) : MDFF_ERR_ARG; }

int mdff_get_thermo(mdff_ctx *c, mdff_thermo *th)
{
    if (!c || !th) return MDFF_ERR_ARG;
    int rc = sync_scal(c);
    if (rc) return rc;
    thermo_from_scal(c, th);
    return MDFF_OK;
}

int mdff_get_ewald(mdff_ctx *c, double *alpha, int *kES, int *nk)
{
    if (!c) return MDFF_ERR_ARG;
    if (alpha) *alpha = c->alpha;
    if (kES) memcpy(kES, c->kES, sizeof(int) * 3);
    if (nk) *nk = c->nk;
    return MDFF_OK;
}

int mdff_init_velocities(mdff_ctx *c)
{
    if (!c) return MDFF_ERR_ARG;
    int n = c->n;
    std::vector<double> vx(n), vy(n), vz(n);
    int rc = mdff_host_init_velocities(&c->p, n, c->mass.data(), vx.data(), vy.data(), vz.data(), 1);
    if (rc) return rc;
    return mdff_set_velocities(c, vx.data(), vy.data(), vz.data());
}

int mdff_check_verlet(mdff_ctx *c)
{
    if (!c) return MDFF_ERR_ARG;
    return build_list(c, 0);
}

// ---------------------------------------------- the check without allocation
// mdff_check_verlet_noalloc: the check and rebuild of mdff_check_verlet with
// every buffer beyond the context's own (the check's block partials and
// finished-block count, the cell grid's five arrays) in `scratch`, one
// array per GPU that lives until the process ends, and no growth of the
// row capacity.  Its address on each device comes from hipGetSymbolAddress
// on the first call there (the array's name is not a device address on the
// host).  Calls that overlap on one GPU share the array and race.
#define NA_MAX_ION (1 << 17)
#define NA_MAX_CELL (1 << 17)
#define NA_OFF_PART 0                                               // double [2][NA_MAX_ION / 256]
#define NA_OFF_CNT (NA_OFF_PART + 16 * (NA_MAX_ION / 256))          // unsigned, 256-byte slot
#define NA_OFF_CCOUNT (NA_OFF_CNT + 256)                            // int [NA_MAX_CELL]
#define NA_OFF_CSTART (NA_OFF_CCOUNT + 4 * NA_MAX_CELL)             // int [NA_MAX_CELL + 1], padded
#define NA_OFF_IONS (NA_OFF_CSTART + 4 * NA_MAX_CELL + 256)         // int [4][NA_MAX_ION]
#define NA_BYTES (NA_OFF_IONS + 16 * NA_MAX_ION)
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

#define NA_MAX_DEV 64
static void *g_scratch_ptr[NA_MAX_DEV];
static std::mutex g_scratch_mu;

static int scratch_address(int device, unsigned char **out)
{
    if (device < 0 || device >= NA_MAX_DEV) return MDFF_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess || hipSetDevice(device) != hipSuccess) return MDFF_ERR_DEVICE;
        void *p = nullptr;
        const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(cur);
        if (err != hipSuccess || !p) return MDFF_ERR_DEVICE;
        g_scratch_ptr[device] = p;
    }
    *out = (unsigned char *)g_scratch_ptr[device];
    return MDFF_OK;
}

int mdff_check_verlet_noalloc(mdff_ctx *c)
{
    if (!c) return MDFF_ERR_ARG;
    const int n = c->n;
    CellGrid g{};
    const bool cells = grid_dims(c, g);
    if (n > NA_MAX_ION || (cells && g.ncell > NA_MAX_CELL)) return MDFF_ERR_SCRATCH;
    unsigned char *base = nullptr;
    int rc = scratch_address(c->device, &base);
    if (rc) return rc;
    double *part = (double *)(base + NA_OFF_PART);
    unsigned *bcnt = (unsigned *)(base + NA_OFF_CNT);
    g.ccount = (int *)(base + NA_OFF_CCOUNT);
    g.cstart = (int *)(base + NA_OFF_CSTART);
    g.cidx = (int *)(base + NA_OFF_IONS);
    g.csort = g.cidx + NA_MAX_ION;
    g.cell = g.csort + NA_MAX_ION;
    g.slot = g.cell + NA_MAX_ION;
    // what must start at zero: the finished-block count and the cell counts
    // (the kernels leave both at zero, a call cut short may not)
    hipLaunchKernelGGL(k_zero, dim3(cdiv((256 + 4 * (size_t)NA_MAX_CELL) / 8, 256)), dim3(256), 0, c->st,
                       (double *)(base + NA_OFF_CNT), (256 + 4 * (size_t)NA_MAX_CELL) / 8);
    hipLaunchKernelGGL(k_frac, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->cd, c->r, c->r + n, c->r + 2 * n, c->s);
    const double csr = c->p.cutshortrange + c->p.skindiff, clr = c->p.cutlongrange + c->p.skindiff;
    const int all = !c->p.lverletL;
    if (all) {
        set_int(c, C_OVERFLOW, 0);
        set_int(c, C_MAXCNT, 0);
        set_int(c, C_REBUILD, c->have_list ? 0 : 1);
    } else {
        if (!c->have_list) set_int(c, C_FIRST, 1);
        hipLaunchKernelGGL(k_vcheck, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->cd, c->s, c->s0, c->p.skindiff,
                           c->ctl, 1, part, bcnt);
    }
    if ((rc = sync_ctl(c))) return rc;
    if (!c->ctl_h[C_REBUILD]) return MDFF_OK;
    launch_list_build(c, g, cells, all, csr, clr);
    if ((rc = sync_ctl(c))) return rc;
    if (c->ctl_h[C_OVERFLOW]) {
        // the rows would need more room than the context holds
        c->have_list = 0;
        return MDFF_ERR_LIST_OVERFLOW;
    }
    c->have_list = 1;
    return MDFF_OK;
}

static int finish_call(mdff_ctx *c, int rc, mdff_thermo *th)
{
    int r2 = sync_scal(c);
    if (r2) return r2;
    if (c->ctl_h[C_NAN]) rc = MDFF_ERR_NAN;
    if (c->have_events) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) c->t_ms[0] = ms;
        if (hipEventElapsedTime(&ms, c->ev[1], c->ev[2]) == hipSuccess) c->t_ms[1] = ms;
        if (hipEventElapsedTime(&ms, c->ev[2], c->ev[3]) == hipSuccess) c->t_ms[2] = ms;
        if (hipEventElapsedTime(&ms, c->ev[3], c->ev[4]) == hipSuccess) c->t_ms[3] = ms;
        if (hipEventElapsedTime(&ms, c->ev[0], c->ev[4]) == hipSuccess) c->t_ms[5] = ms;
    }
    if (th) thermo_from_scal(c, th);
    return rc;
}

int mdff_engforce(mdff_ctx *c, int istep, mdff_thermo *th)
{
    if (!c) return MDFF_ERR_ARG;
    int rc = engforce_impl(c, istep);
    if (rc && rc != MDFF_ERR_SCF) return rc;
    return finish_call(c, rc, th);
}

int mdff_pair_short(mdff_ctx *c, double *u, double *pvir, double tau[9])
{
    if (!c) return MDFF_ERR_ARG;
    int n = c->n;
    hipLaunchKernelGGL(k_frac, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->cd, c->r, c->r + n, c->r + 2 * n, c->s);
    hipLaunchKernelGGL(k_zero, dim3(cdiv(3 * n, 256)), dim3(256), 0, c->st, c->f_sr, (size_t)3 * n);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(14), 0, c->st, c->scal + S_U_BH, (size_t)14);
    short_range(c);
    HIPCHK(hipMemcpyAsync(c->f, c->f_sr, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, c->st));
    int rc = sync_scal(c);
    if (rc) return rc;
    mdff_thermo t;
    thermo_from_scal(c, &t);
    const double(*tt)[3] = c->p.lnmlj ? t.tau_lj : t.tau_bhmftd;
    if (u) *u = c->p.lnmlj ? t.u_lj : t.u_bhmftd;
    if (pvir) *pvir = c->p.lnmlj ? t.pvir_lj : t.pvir_bhmftd;
    if (tau)
        for (int i = 0; i < 9; i++) tau[i] = tt[i / 3][i % 3];
    return MDFF_OK;
}

int mdff_get_dipoles_scf(mdff_ctx *c, int istep, double *mu, double *upol)
{
    if (!c) return MDFF_ERR_ARG;
    int n = c->n, rc = MDFF_OK;
    hipLaunchKernelGGL(k_frac, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->cd, c->r, c->r + n, c->r + 2 * n, c->s);
    if (c->lpol) {
        int lo = lo_of(c);
        const bool pol = c->npol > 0;
        launch_eikr(c, 0, lo, nullptr, pol);
        if (pol && lo && scf_uses_fp64(c)) launch_eikr(c, 1, 0, nullptr);
        if ((rc = build_t2(c)) != MDFF_OK) return rc;
        static_field(c, lo);
        rc = c->p.algo_pim == MDFF_PIM_SCFKO ? scf_loop_ko(c, istep) : (lo ? scf_mixed(c, istep) : scf_loop(c, istep));
        if (rc && rc != MDFF_ERR_SCF) return rc;
        c->ldip_sticky = 1;
        hipLaunchKernelGGL(k_axpy3, dim3(cdiv(3 * n, 256)), dim3(256), 0, c->st, 3 * n, c->mu_static, c->mu_ind, c->mu_tot);
    } else {
        HIPCHK(hipMemcpyAsync(c->mu_tot, c->mu_static, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, c->st));
    }
    int r2 = sync_scal(c);
    if (r2) return r2;
    if (upol) *upol = c->lpol ? c->scal_h[S_UPOL] * c->u.coul_unit : 0.0;
    if (mu) {
        r2 = get_aos3(c, c->mu_tot, mu);
        if (r2) return r2;
    }
    return rc;
}

int mdff_multipole_es(mdff_ctx *c, const double *qh, const double *muh, const double *theta, int flags, double *u,
                      double *pvir, double tau[9], double *efh, double *efgh, double *fx, double *fy, double *fz)
{
    (void)theta;
    if (!c || !qh || !muh) return MDFF_ERR_ARG;
    if (!c->p.lcoulombic) return MDFF_ERR_CONFIG;
    int n = c->n;
    const int lq = flags & MDFF_ES_LQ, ld = flags & MDFF_ES_LD, ldamp = flags & MDFF_ES_LDAMP;
    const int do_f = flags & MDFF_ES_FORCES, do_s = flags & MDFF_ES_STRESS, do_ef = flags & MDFF_ES_EF,
              do_efg = flags & MDFF_ES_EFG, do_dir = flags & MDFF_ES_DIR, do_rec = flags & MDFF_ES_REC;
    // terms masked by lq/ld; the self terms and the reciprocal charge force
    // use the arrays as passed (ewald.c:193-226, :984)
    std::vector<double> qm(n), qr(n), mum(3 * n), mur(3 * n);
    for (int i = 0; i < n; i++) {
        qr[i] = qh[i];
        qm[i] = lq ? qh[i] : 0.0;
        for (int k = 0; k < 3; k++) {
            mur[(size_t)k * n + i] = muh[3 * i + k];
            mum[(size_t)k * n + i] = ld ? muh[3 * i + k] : 0.0;
        }
    }
    double *d_qm = c->tmp_q, *d_qr = c->tmp_qraw, *d_mum = c->tmp_mu, *d_mur = c->tmp_mur;
    HIPCHK(hipMemcpyAsync(d_qm, qm.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(d_qr, qr.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(d_mum, mum.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(d_mur, mur.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL(k_frac, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->cd, c->r, c->r + n, c->r + 2 * n, c->s);
    hipLaunchKernelGGL(k_zero, dim3(cdiv(3 * n, 256)), dim3(256), 0, c->st, c->ef_dir, (size_t)3 * n);
    hipLaunchKernelGGL(k_zero, dim3(cdiv(9 * n, 256)), dim3(256), 0, c->st, c->efg_dir, (size_t)9 * n);
    hipLaunchKernelGGL(k_zero, dim3(cdiv(3 * n, 256)), dim3(256), 0, c->st, c->f_dir, (size_t)3 * n);
    hipLaunchKernelGGL(k_zero, dim3(cdiv(7 * n, 256)), dim3(256), 0, c->st, c->uw_dir, (size_t)7 * n);
    hipLaunchKernelGGL(k_zero, dim3(cdiv(3 * n, 256)), dim3(256), 0, c->st, c->ef_rec, (size_t)3 * n);
    hipLaunchKernelGGL(k_zero, dim3(cdiv(9 * n, 256)), dim3(256), 0, c->st, c->efg_rec, (size_t)9 * n);
    hipLaunchKernelGGL(k_zero, dim3(cdiv(3 * n, 256)), dim3(256), 0, c->st, c->f_rec, (size_t)3 * n);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, c->st, c->scal + S_U_DIR, (size_t)14);
    if (do_dir) {
        DirArgs a = dir_args(c, d_qm, d_mum, ldamp ? 1 : 0, 0, nullptr);
        a.efg = c->efg_dir;
        launch_dir(c, a, 0, do_ef, do_efg, do_efg || do_f, do_f);
        colsum(c, c->uw_dir, n, n, 7, c->scal + S_U_DIR);
    }
    if (do_rec) {
        launch_eikr(c, 0, 0, nullptr);
        int C = ld ? 4 : 1;
        int NCH = (do_ef || (do_f && ld)) ? 9 : 3;
        rec_part(c, 0, 0, C, NCH, d_qm, d_mum, d_qr, d_mum, do_ef, do_f, do_s, c->efg_rec, nullptr);
        colsum(c, c->ek, (int)c->nkf, (int)c->nkf, 7, c->scal + S_U_REC);
    }
    hipLaunchKernelGGL(k_self_terms, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, d_qr, d_mur, c->selfbuf);
    colsum(c, c->selfbuf, n, n, 2, c->scal + S_QSQ);
    std::vector<double> hdir(3 * n), hrec(3 * n), gdir(9 * n), grec(9 * n), fdir(3 * n), frec(3 * n);
    HIPCHK(hipMemcpyAsync(hdir.data(), c->ef_dir, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(hrec.data(), c->ef_rec, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(gdir.data(), c->efg_dir, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(grec.data(), c->efg_rec, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(fdir.data(), c->f_dir, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(frec.data(), c->f_rec, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->st));
    int rc = sync_scal(c);
    if (rc) return rc;
    const double *S = c->scal_h;
    double u_self = -c->selfa * S[S_QSQ] - c->selfa2 * S[S_MUSQ];
    double udir = do_dir ? S[S_U_DIR] : 0.0, urec = do_rec ? S[S_U_REC] * c->ttpiV : 0.0;
    if (u) *u = (udir + urec + u_self) * c->u.coul_unit;
    double iv = c->cell.inveOmegaPU;
    if (do_s && tau) {
        const double *wd = S + S_W_DIR, *wr = S + S_W_REC;
        int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
        for (int k = 0; k < 9; k++) {
            double t = 0.0;
            if (do_dir && do_f) t += 0.5 * wd[map[k]] * iv;
            if (do_rec) t += wr[map[k]] * c->ttpiV * iv;
            tau[k] = t * c->u.coul_unit;
        }
    }
    if (pvir && tau) *pvir = (tau[0] + tau[4] + tau[8]) * (1.0 / 3.0);
    for (int i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) {
            double v = 2.0 * c->selfa2 * muh[3 * i + a];
            if (do_dir && do_ef) v += hdir[(size_t)a * n + i];
            if (do_rec && do_ef) v += hrec[(size_t)a * n + i];
            if (efh) efh[3 * i + a] = v;
            for (int b = 0; b < 3; b++) {
                double g = (a == b) ? -2.0 * c->selfa2 * qh[i] : 0.0;
                if (do_dir) g += gdir[(size_t)(3 * a + b) * n + i];
                if (do_rec && do_ef) g += grec[(size_t)(3 * a + b) * n + i];
                if (efgh) efgh[9 * i + 3 * a + b] = g;
            }
        }
        if (do_f) {
            double fxi = 0, fyi = 0, fzi = 0;
            if (do_dir) { fxi += fdir[i]; fyi += fdir[n + i]; fzi += fdir[2 * n + i]; }
            if (do_rec) { fxi += frec[i]; fyi += frec[n + i]; fzi += frec[2 * n + i]; }
            if (fx) fx[i] += fxi * c->u.coul_unit;
            if (fy) fy[i] += fyi * c->u.coul_unit;
            if (fz) fz[i] += fzi * c->u.coul_unit;
        }
    }
    return MDFF_OK;
}

int mdff_md_init(mdff_ctx *c, mdff_thermo *th)
{
    if (!c) return MDFF_ERR_ARG;
    int n = c->n;
    hipLaunchKernelGGL(k_ke, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->v, c->massd, c->ke);
    kin_reduce(c);
    int rc = engforce_impl(c, 0);
    if (rc && rc != MDFF_ERR_SCF) return rc;
    rc = finish_call(c, rc, th);
    if (th) th->istep = 0;
    return rc;
}

int mdff_md_step(mdff_ctx *c, int istep, mdff_thermo *th)
{
    if (!c) return MDFF_ERR_ARG;
    c->kin_fresh = 0;   // velocities may have been set since the last call
    int rc = md_step_impl(c, istep);
    if (rc && rc != MDFF_ERR_SCF) return rc;
    rc = finish_call(c, rc, th);
    if (th) th->istep = istep;
    return rc;
}

int mdff_propagate(mdff_ctx *c, int istep, mdff_thermo *th)
{
    if (!c) return MDFF_ERR_ARG;
    c->kin_fresh = 0;   // velocities may have been set since the last call
    int rc = prop_step(c, istep);
    if (rc && rc != MDFF_ERR_SCF) return rc;
    rc = finish_call(c, rc, th);
    if (th) th->istep = istep;
    return rc;
}

int mdff_md_run(mdff_ctx *c, int first_step, int nsteps, mdff_thermo *th)
{
    if (!c) return MDFF_ERR_ARG;
    int worst = MDFF_OK;
    c->kin_fresh = 0;   // velocities may have been set since the last call
    for (int k = 0; k < nsteps; k++) {
        int rc = md_step_impl(c, first_step + k);
        if (rc == MDFF_ERR_SCF) worst = rc;
        else if (rc) return rc;
    }
    int rc = finish_call(c, worst, th);
    if (th) th->istep = first_step + nsteps - 1;
    return rc;
}

int mdff_time_kernel(mdff_ctx *c, int which, int lo, int iters, double *ms)
{
    if (which == 7) {
        // forced rebuilds of the neighbour list at the current positions by
        // the method in use (the next check measures displacements from them)
        if (!c || !ms || iters < 1) return MDFF_ERR_ARG;
        int n = c->n;
        if (!c->cells_set && setup_cells(c)) return MDFF_ERR_NOMEM;
        hipLaunchKernelGGL(k_frac, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->cd, c->r, c->r + n, c->r + 2 * n,
                           c->s);
        const double csr = c->p.cutshortrange + c->p.skindiff, clr = c->p.cutlongrange + c->p.skindiff;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return MDFF_ERR_DEVICE;
        set_int(c, C_REBUILD, 1);
        launch_list_build(c, c->cg, c->cells_on, !c->p.lverletL, csr, clr);
        HIPCHK(hipEventRecord(e0, c->st));
        for (int k = 0; k < iters; k++) launch_list_build(c, c->cg, c->cells_on, !c->p.lverletL, csr, clr);
        HIPCHK(hipEventRecord(e1, c->st));
        float t = 0.0f;
        int rc = (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&t, e0, e1) == hipSuccess)
                     ? MDFF_OK : MDFF_ERR_DEVICE;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        set_int(c, C_REBUILD, 0);
        *ms = t / iters;
        return rc;
    }
    if (!c || !ms || iters < 1 || !c->p.lcoulombic || !(lo == 0 || lo == 1 || lo == 3)) return MDFF_ERR_ARG;
    if (lo && !lo_of(c)) return MDFF_ERR_ARG;
    // lo = 3: FP32 tables with the split-FP16 matrix-core mode
    struct ModeGuard {
        mdff_ctx *c; int keep;
        ~ModeGuard() { c->mf_lo = keep; }
    } guard{c, c->mf_lo};
    if (lo == 1) c->mf_lo = KM_F32;
    else if (lo == 3) { c->mf_lo = KM_F16X3; lo = 1; }
    int n = c->n;
    int polar = (which == 0 || which == 1 || which == 2);
    if (polar && c->npol == 0) return MDFF_ERR_ARG;
    // state for the kernels: tables, a dipole field and W from one full pass
    hipLaunchKernelGGL(k_frac, dim3(cdiv(n, 256)), dim3(256), 0, c->st, n, c->cd, c->r, c->r + n, c->r + 2 * n, c->s);
    launch_eikr(c, polar, lo, nullptr);
    const double *mu = polar ? c->mu_ind : c->mu_tot;
    int C = polar ? 3 : 4, NCH = polar ? 3 : 9;
    rec_part(c, polar, lo, C, NCH, polar ? c->qz : c->q, mu, polar ? c->qz : c->q, mu, 1, !polar, 0, nullptr, nullptr);
    int nsel = polar ? c->npol : n;
    SkArgs a;
    a.nsel = nsel; a.Kx = c->Kx; a.Ky = c->Ky; a.Kz = c->Kz; a.P = c->P;
    a.tx = c->tab3[lo][polar][0]; a.ty = c->tab3[lo][polar][1]; a.tz = c->tab3[lo][polar][2];
    a.idx = polar ? c->pidx : nullptr; a.q = polar ? c->qz : c->q; a.mu = mu; a.n = n;
    a.chunk = polar ? c->chunk_p : c->chunk; a.part = c->kpart; a.active = nullptr;
    int nchunk = polar ? c->nchunk_p : c->nchunk;
    BackArgs b;
    b.nsel = nsel; b.Kx = c->Kx; b.Ky = c->Ky; b.Kz = c->Kz; b.P = c->P;
    b.tx = a.tx; b.ty = a.ty; b.tz = a.tz; b.W = c->W; b.rows_per_split = c->rows_per_split; b.part = c->bpart;
    b.active = nullptr;
    DirArgs d = dir_args(c, polar ? c->qz : c->q, mu, polar ? 0 : c->ldmp, polar ? 1 : 0, nullptr);
    if (which == 2 && build_t2(c, true) != MDFF_OK) return MDFF_ERR_NOMEM;
    auto once = [&]() {
        int dummy = 0;
        switch (which) {
        case 0: case 3: {
            // as a force call launches it: the split-FP16 sum leaves its
            // partials to the reduction (and folds the SCF's dipole channels)
            c->psum_defer = !c->no_fuse;
            c->psum_ch = NCH == 9 ? 6 : 3;
            const bool done = launch_fwd_mfma(c, polar, lo, C, a.q, mu, nullptr, &dummy);
            c->psum_defer = false;
            if (done) break;
            if (lo) { if (C == 3) launch_fwd_t<float, 3>(c, a, nchunk); else launch_fwd_t<float, 4>(c, a, nchunk); }
            else { if (C == 3) launch_fwd_t<double, 3>(c, a, nchunk); else launch_fwd_t<double, 4>(c, a, nchunk); }
            break;
        }
        case 1: case 4:
            if (launch_back_mfma(c, polar, lo, NCH, nullptr, &dummy)) break;
            if (lo) { if (NCH == 3) launch_back_t<float, 3>(c, b, nsel); else launch_back_t<float, 9>(c, b, nsel); }
            else { if (NCH == 3) launch_back_t<double, 3>(c, b, nsel); else launch_back_t<double, 9>(c, b, nsel); }
            break;
        case 2:
            if (c->t2_ok) {
                unsigned blocks = cdiv((size_t)n * 16, 256);
                if (lo)
                    hipLaunchKernelGGL((k_t2_field<float, 16>), dim3(blocks), dim3(256), 0, c->st, n, c->capP,
                                       c->npol, c->pinv, c->cntp, c->idxp, c->t2f, mu, c->ef_dir, nullptr);
                else
                    hipLaunchKernelGGL((k_t2_field<double, 16>), dim3(blocks), dim3(256), 0, c->st, n, c->capP,
                                       c->npol, c->pinv, c->cntp, c->idxp, c->t2, mu, c->ef_dir, nullptr);
            } else {
                launch_dir(c, d, lo, 1, 0, 0, 0);
            }
            break;
        case 5: short_range(c); break;
        case 6: launch_dir(c, d, lo, 0, 0, 0, 1); break;
        default: break;
        }
    };
    once();
    c->fwd_part = 0;
    // the k-space GEMM kernels are timed alone (operand builds and the
    // partial-sum pass excluded): `iters` launches of the GEMM back to back
    // between one pair of events, inside one pass of the sequence
    std::vector<hipEvent_t> pe;
    const bool gemm = which == 0 || which == 1 || which == 3 || which == 4;
    if (gemm) {
        pe.resize(2);
        bool ok = hipEventCreate(&pe[0]) == hipSuccess;
        ok = hipEventCreate(&pe[1]) == hipSuccess && ok;
        if (ok) { c->probe_ev = pe.data(); c->probe_k = 0; c->probe_max = 1; c->probe_rep = iters; }
        else pe.clear();
    }
    bool timed = false;
    if (c->probe_ev) {
        once();
        // a sum that took another path (vector ALU) set no probe events
        const bool hit = c->probe_k == 1;
        c->probe_ev = nullptr;
        c->probe_rep = 1;
        float tk = 0.0f;
        if (hit && hipEventSynchronize(pe[1]) == hipSuccess && hipEventElapsedTime(&tk, pe[0], pe[1]) == hipSuccess) {
            *ms = (double)tk / iters;
            timed = true;
        }
    }
    if (!timed) {
        HIPCHK(hipEventRecord(c->ev[6], c->st));
        for (int k = 0; k < iters; k++) once();
        HIPCHK(hipEventRecord(c->ev[7], c->st));
        HIPCHK(hipEventSynchronize(c->ev[7]));
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, c->ev[6], c->ev[7]));
        *ms = (double)t / iters;
    }
    c->probe_ev = nullptr;
    c->probe_k = c->probe_max = 0;
    c->probe_rep = 1;
    for (auto e : pe)
        if (e) (void)hipEventDestroy(e);
    c->fwd_part = 0;
    return MDFF_OK;
}

int mdff_get_counters(mdff_ctx *c, long long out[4])
{
    if (!c || !out) return MDFF_ERR_ARG;
    out[0] = c->n_force;
    out[1] = c->n_scf_fp64;
    out[2] = c->n_scf_inner;
    out[3] = c->npol;
    return MDFF_OK;
}

int mdff_set_scf_solver(mdff_ctx *c, int solver, int inner)
{
    if (!c || (solver != 0 && solver != 2) || inner < 0 || inner > 64) return MDFF_ERR_ARG;
    if (lo_of(c)) c->scf_solver = solver;
    else c->fp64_cg = solver == 2;
    c->cg_first = inner;
    c->hint_cg = 0;
    return MDFF_OK;
}

int mdff_set_list_method(mdff_ctx *c, int method)
{
    if (!c || method < 0 || method > 2) return MDFF_ERR_ARG;
    if (method == c->list_method) return MDFF_OK;
    c->list_method = method;
    free_cells(c);
    return MDFF_OK;
}

int mdff_list_info(mdff_ctx *c, int out[8])
{
    if (!c || !out) return MDFF_ERR_ARG;
    out[0] = c->cells_on;
    for (int d = 0; d < 3; d++) {
        out[1 + d] = c->cells_on ? c->cg.nc[d] : 0;
        out[4 + d] = c->cells_on ? c->cg.reach[d] : 0;
    }
    out[7] = c->cap;
    return MDFF_OK;
}

int mdff_set_scf_check(mdff_ctx *c, int every)
{
    if (!c || every < 0) return MDFF_ERR_ARG;
    c->scf_check = every;
    return MDFF_OK;
}

int mdff_set_scf_refine(mdff_ctx *c, int on)
{
    if (!c || on < 0 || on > 1) return MDFF_ERR_ARG;
    c->scf_refine = on;
    c->hint_ref = 0;
    return MDFF_OK;
}

int mdff_get_timing(mdff_ctx *c, double t[6])
{
    if (!c || !t) return MDFF_ERR_ARG;
    for (int i = 0; i < 6; i++) t[i] = c->t_ms[i];
    return MDFF_OK;
}

}  // extern "C"
