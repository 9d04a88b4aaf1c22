/*
 * mdff_dropin.c -- same-name drop-ins for mdff20's force path
 * (declarations and contract in include/mdff_dropin.h).
 *
 * Build: compiled against the reference's headers, which define its globals
 * (compile with -fcommon as the reference itself needs), and linked with
 * libmdff_gpu.so into the reference's executable in place of the functions
 * it replaces, e.g.
 *   cc -c mdff_dropin.c -I<mdff20>/src -I<repo>/include -fcommon
 * Each entry point maps the reference's global state onto the handle API of
 * mdff_gpu.h; no computation happens here.
 */
#include <math.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* the reference's globals (its own headers, from its source tree) */
#include "bhmftd.h"
#include "cell.h"
#include "config.h"
#include "constants.h"
#include "coulombic.h"
#include "ewald.h"
#include "field.h"
#include "global.h"
#include "integration.h"
#include "inverse_matrix.h"
#include "io.h"
#include "kinetic.h"
#include "md.h"
#include "nmlj.h"
#include "pim.h"
#include "radial_distribution.h"
#include "thermo.h"
#include "timing.h"
#include "verlet.h"

#include "mdff_dropin.h"
#include "mdff_gpu.h"
#include "mdff_host.h"
#include "mdff_rdf.h"

static mdff_ctx *g_ctx = NULL;
static int g_nion = 0;
static int g_atexit = 0;

/* The reference draws its initial velocities from the C library's rand()
 * (kinetic.c, rand.c), and the HIP runtime may draw from the same generator.
 * Every call into the device library therefore runs on a private generator
 * state; the reference's sequence is left where it was. */
static char g_rng_state[256];
static int g_rng_ready = 0;

static char *rng_enter(void)
{
    if (!g_rng_ready) {
        g_rng_ready = 1;
        return initstate(1u, g_rng_state, sizeof(g_rng_state));
    }
    return setstate(g_rng_state);
}

static void rng_leave(char *prev) { setstate(prev); }

static void fail(const char *where, int rc)
{
    fprintf(stderr, "mdff drop-in: %s failed with status %d\n", where, rc);
    exit(-1);
}

/* mdff_params from the parsed globals of global.c, field.c, bhmftd.c,
 * nmlj.c, pim.c and config.c; alphaES/kES already hold what set_autoES chose */
static void params_from_globals(mdff_params *p)
{
    mdff_params_default(p);
    p->ntype = ntype;
    p->Fposff = Fposff;
    p->lverletL = lverletL;
    p->skindiff = skindiff;
    p->lstatic = lstatic;
    p->lrdf = lrdf;
    p->lpstress = lpstress;
    p->lreduced = lreduced;
    /* the cutoffs the force routines use: lrcutsq is set before set_autoES
     * overwrites cutlongrange (field.c:329-330) */
    p->cutshortrange = srcutsq > 0.0 ? sqrt(srcutsq) : cutshortrange;
    p->cutlongrange = lrcutsq > 0.0 ? sqrt(lrcutsq) : cutlongrange;
    /* the integrator of md.c's globals (check_md has converted them to
     * internal units, md.c:120-141; the context takes control-file units) for
     * prop_agate / run_md; a static run only evaluates forces */
    if (lstatic) {
        p->integrator = MDFF_NVE_VV;
        p->nequil = 0;
    } else {
        p->integrator = egrator == 0 ? MDFF_NVE_LF : (egrator == 2 ? MDFF_NVT_NHCN : MDFF_NVE_VV);
        p->dt = dt / time_unit;
        p->temp = temp / boltz_unit;
        p->tauTberendsen = tauTberendsen / time_unit;
        p->timesca_thermo = timesca_thermo / time_unit;
        p->nhc_n = nhc_n;
        p->nhc_mults = nhc_mults;
        p->nhc_yosh_order = nhc_yosh_order;
        p->npas = npas;
        p->nequil = nequil;
        p->nequilT = nequilT;
    }
    p->lnmlj = lnmlj;
    p->lbhmft = lbhmft;
    p->lbhmftd = lbhmftd;
    p->lcoulombic = lcoulombic;
    p->lsymmetric = lsymmetric;
    for (int t = 0; t < NTYPEMAX && t < MDFF_NTYPEMAX; t++) {
        p->massit[t] = massit[t];
        p->qit[t] = qit[t];
        p->lpolar[t] = lpolar[t];
        for (int i = 0; i < 3; i++) {
            p->dipit[t][i] = dipit[t][i];
            for (int j = 0; j < 3; j++) {
                p->quadit[t][i][j] = quadit[t][i][j];
                p->polit[t][i][j] = polit[t][i][j];
            }
        }
        for (int i = 0; i < NTYPEMAX && i < MDFF_NTYPEMAX; i++) {
            p->Abhmftd[t][i] = Abhmftd[t][i];
            p->Bbhmftd[t][i] = Bbhmftd[t][i];
            p->Cbhmftd[t][i] = Cbhmftd[t][i];
            p->Dbhmftd[t][i] = Dbhmftd[t][i];
            p->BDbhmftd[t][i] = BDbhmftd[t][i];
            p->sigmalj[t][i] = sigmalj[t][i];
            p->epslj[t][i] = epslj[t][i];
            p->plj[t][i] = plj[t][i];
            p->qlj[t][i] = qlj[t][i];
            for (int k = 0; k < NTYPEMAX && k < MDFF_NTYPEMAX; k++) {
                p->lpoldamping[t][i][k] = lpoldamping[t][i][k];
                p->pol_damp_b[t][i][k] = pol_damp_b[t][i][k];
                p->pol_damp_c[t][i][k] = pol_damp_c[t][i][k];
                p->pol_damp_k[t][i][k] = pol_damp_k[t][i][k];
            }
        }
    }
    p->trunc = trunctype;
    p->alphaES = alphaES;
    for (int k = 0; k < 3; k++) p->kES[k] = kES[k];
    p->epsw = epsw;
    p->lautoES = 0;
    p->algo_pim = algo_pim;
    p->algo_extrapolate_dipole = algo_extrapolate_dipole;
    p->extrapolate_order = extrapolate_order;
    p->min_scf_pim_iter = min_scf_pim_iter;
    p->max_scf_pim_iter = max_scf_pim_iter;
    p->conv_tol_ind = conv_tol_ind;
    p->omegakO = omegakO;
}

int mdff_dropin_params(mdff_params *p)
{
    if (!p) return MDFF_ERR_ARG;
    params_from_globals(p);
    return MDFF_OK;
}

/* the histogram handle of get_gr and the sizes it was made for */
static mdff_rdf *g_rdf = NULL;
static int g_rdf_n = 0, g_rdf_nt = 0, g_rdf_nb = 0;
static double g_rdf_resg = 0.0, g_rdf_cut = 0.0;
static int *g_rdf_type = NULL;

void mdff_dropin_release(void)
{
    char *rs = rng_enter();
    if (g_ctx) mdff_destroy(g_ctx);
    if (g_rdf) mdff_rdf_destroy(g_rdf);
    rng_leave(rs);
    g_ctx = NULL;
    g_nion = 0;
    g_rdf = NULL;
    free(g_rdf_type);
    g_rdf_type = NULL;
}

/* the device context for the current system; positions uploaded */
static mdff_ctx *ctx_sync(void)
{
    if (!g_ctx || g_nion != nion) {
        mdff_dropin_release();
        mdff_params p;
        params_from_globals(&p);
        double A9[9];
        for (int i = 0; i < 9; i++) A9[i] = simuCell.A[i / 3][i % 3];
        const char *pr = getenv("MDFF_PRECISION");
        int prec = (pr && !strcmp(pr, "mixed")) ? MDFF_MIXED : MDFF_FP64;
        const char *dv = getenv("MDFF_DEVICE");
        int dev = dv ? atoi(dv) : 0;
        int rc = mdff_create(&g_ctx, &p, nion, typia, A9, rx, ry, rz, prec, dev);
        if (rc) fail("mdff_create", rc);
        g_nion = nion;
        if (!g_atexit) {
            atexit(mdff_dropin_release);
            g_atexit = 1;
        }
        return g_ctx;
    }
    int rc = mdff_set_positions(g_ctx, rx, ry, rz);
    if (rc) fail("mdff_set_positions", rc);
    return g_ctx;
}

static void copy_tau(double dst[3][3], const double src[3][3])
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) dst[i][j] = src[i][j];
}

/* the induced dipoles into the reference's history slot 0, after the shift
 * the predictor of extrapolate_dipole_aspc performs (pim.c:393-398) */
static void store_induced(mdff_ctx *c)
{
    if (!lpim || !dipia_ind) return;
    double *mu = (double *)malloc(sizeof(double) * 3 * (size_t)nion);
    if (!mu) fail("malloc", MDFF_ERR_NOMEM);
    int rc = mdff_get_dipoles(c, mu);
    if (rc) fail("mdff_get_dipoles", rc);
    for (int ia = 0; ia < nion; ia++) {
        for (int k = extrapolate_order; k > 0; k--)
            for (int i = 0; i < 3; i++) dipia_ind[ia][k][i] = dipia_ind[ia][k - 1][i];
        for (int i = 0; i < 3; i++) dipia_ind[ia][0][i] = mu[3 * ia + i] - dipia[ia][i];
    }
    free(mu);
}

/* the globals engforce sets (field.c:352-407) from one force call */
static void force_globals(mdff_ctx *c, const mdff_thermo *t, int count_scf)
{
    mdff_thermo th = *t;
    if (lnmlj) {
        u_lj = th.u_lj;
        pvir_lj = th.pvir_lj;
        copy_tau(tau_lj, th.tau_lj);
    }
    if (lbhmft || lbhmftd) {
        u_bhmftd = th.u_bhmftd;
        pvir_bhmftd = th.pvir_bhmftd;
        copy_tau(tau_bhmftd, th.tau_bhmftd);
    }
    if (lcoulombic) {
        u_coul = th.u_coul;
        pvir_coul = th.pvir_coul;
        copy_tau(tau_coul, th.tau_coul);
        u_pol = th.u_pol;
        if (lpol) {
            ldip = true;
            /* the dipole Ewald calls the reference's SCF would have made */
            if (count_scf) {
                cc_multipole_ES_dir_pim += th.n_scf + 1;
                cc_multipole_ES_rec_pim += th.n_scf + 1;
            }
            store_induced(c);
        }
        for (int it = 0; it < ntype; it++)
            for (int j = 0; j < 3; j++)
                if (dipit[it][j] != 0.0) ldip = true;
    }
}

static void engforce_impl(void)
{
    mdff_ctx *c = ctx_sync();
    mdff_thermo th;
    int rc = mdff_engforce(c, istep, &th);
    if (rc && rc != MDFF_ERR_SCF) fail("mdff_engforce", rc);
    rc = mdff_get_forces(c, fx, fy, fz);
    if (rc) fail("mdff_get_forces", rc);
    force_globals(c, &th, 1);
}

/* positions, velocities and forces of the device into the globals */
static void state_down(mdff_ctx *c)
{
    int rc = mdff_get_positions(c, rx, ry, rz);
    if (!rc) rc = mdff_get_velocities(c, vx, vy, vz);
    if (!rc) rc = mdff_get_forces(c, fx, fy, fz);
    if (rc) fail("state download", rc);
}

/* the globals a propagator sets beside the force call's: e_kin, and e_nvt
 * of the chain (integration.c:22-294) */
static void step_globals(mdff_ctx *c, const mdff_thermo *th, int count_scf)
{
    force_globals(c, th, count_scf);
    e_kin = th->e_kin;
    e_nvt = th->e_nvt;
    updatevl = th->verlet_updates;
}

static void prop_agate_impl(int step)
{
    mdff_ctx *c = ctx_sync();
    /* the reference's loop may have rescaled the velocities since */
    int rc = mdff_set_velocities(c, vx, vy, vz);
    if (rc) fail("mdff_set_velocities", rc);
    mdff_thermo th;
    rc = mdff_propagate(c, step, &th);
    if (rc && rc != MDFF_ERR_SCF) fail("mdff_propagate", rc);
    state_down(c);
    step_globals(c, &th, 1);
}

/* CONTFF / TRAJFF records of the globals in write_config's format
 * (config.c:291-345) by the library's writer, which opens the file with a
 * proper mode string (write_config passes the address of one char as the
 * mode, whose undefined tail makes some of its opens fail) */
static void config_out(const char *file, const char *mode)
{
    mdff_config cfg;
    memset(&cfg, 0, sizeof(cfg));
    cfg.nion = nion;
    cfg.ntype = ntype;
    snprintf(cfg.name, sizeof(cfg.name), "%s", configname);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) cfg.A[i][j] = simuCell.A[i][j];
    for (int t = 0; t < ntype && t < MDFF_NTYPEMAX; t++) {
        snprintf(cfg.atypit[t], sizeof(cfg.atypit[t]), "%s", atypit[t]);
        cfg.nionit[t] = nionit[t];
    }
    cfg.typia = typia;
    cfg.rx = rx; cfg.ry = ry; cfg.rz = rz;
    cfg.vx = vx; cfg.vy = vy; cfg.vz = vz;
    cfg.fx = fx; cfg.fy = fy; cfg.fz = fz;
    int rc = mdff_write_config(file, mode, &cfg);
    if (rc) {
        pError("opening file :");
        printf(" %s\n", file);
    }
}

/* whether run_md writes stdout, OSZIFF or CONTFF after step s
 * (md.c:250-271) */
static int output_step(int s)
{
    return iopnode(s, npas, nprint) || iopnode(s, npas, fprint) || iopnode(s, npas, cprint);
}

/* the TRAJFF records held in device memory, written in step order through
 * the globals (write_config's format), then forgotten; the globals are
 * left holding the last frame */
static void traj_flush(mdff_ctx *c)
{
    int info[3];
    int rc = mdff_traj_info(c, info);
    if (rc) fail("mdff_traj_info", rc);
    if (info[2] > 0) fail("device trajectory (frames dropped)", MDFF_ERR_ARG);
    double *r[3] = {rx, ry, rz}, *v[3] = {vx, vy, vz}, *f[3] = {fx, fy, fz};
    for (int k = 0; k < info[0]; k++) {
        if ((rc = mdff_traj_frame(c, k, NULL, r, v, f))) fail("mdff_traj_frame", rc);
        config_out("TRAJFF", "a");
    }
    if ((rc = mdff_traj_clear(c))) fail("mdff_traj_clear", rc);
}

static void run_md_impl(void)
{
    mdff_ctx *c = ctx_sync();
    int rc = mdff_set_velocities(c, vx, vy, vz);
    if (rc) fail("mdff_set_velocities", rc);
    /* e_kin and the force call at t=0 (md.c:196-209) */
    mdff_thermo th;
    rc = mdff_md_init(c, &th);
    if (rc && rc != MDFF_ERR_SCF) fail("mdff_md_init", rc);
    rc = mdff_get_forces(c, fx, fy, fz);
    if (rc) fail("mdff_get_forces", rc);
    step_globals(c, &th, 1);
    long long k0[4];
    if ((rc = mdff_get_counters(c, k0))) fail("mdff_get_counters", rc);
    if (ionode) {
        SEPARATOR;
        printf("properties at t=0\n");
        SEPARATOR;
        putchar('\n');
    }
    info_thermo(0, NULL);
    if (ionode && npas > 0) {
        SEPARATOR;
        printf("starting main MD loop\n");
        SEPARATOR;
        putchar('\n');
    }
    FILE *fpOSZIFF = fopen("OSZIFF", "w");
    if (NULL == fpOSZIFF) {
        pError("opening OSZIFF file");
        exit(-1);
    }
    /* TRAJFF records (md.c:275-277) stay in device memory (at most 1 GiB of
     * frames) until the buffer is full or the run ends */
    int tcap = 0, tpend = 0;
    if (tprint > 0) {
        int nt = 0;
        for (int s = 1; s <= npas; s++) nt += iopsnode(s, npas, tprint, tstart) ? 1 : 0;
        const long long fit = (1LL << 30) / (72LL * nion);
        tcap = nt < fit ? nt : (int)(fit > 0 ? fit : 1);
        if (tcap > 0 && (rc = mdff_set_trajectory(c, tprint, tstart, npas, tcap))) fail("mdff_set_trajectory", rc);
    }
    /* the loop of md.c:228-279 on the device: the steps between two outputs
     * run as one mdff_md_run (prop_agate, Verlet check, rescale, the TRAJFF
     * frame copy), and the state comes to the host only where something is
     * written */
    statime(0);
    for (int first = 1; first <= npas;) {
        int last = first, pend = tpend;
        for (;;) {
            pend += (tcap > 0 && iopsnode(last, npas, tprint, tstart)) ? 1 : 0;
            if (last >= npas || output_step(last) || (tcap > 0 && pend == tcap)) break;
            last++;
        }
        rc = mdff_md_run(c, first, last - first + 1, &th);
        if (rc && rc != MDFF_ERR_SCF) fail("mdff_md_run", rc);
        tpend = pend;
        istep = last;
        step_globals(c, &th, 0);
        if (iopnode(istep, npas, nprint)) {
            statime(1);
            mestime(&mdstepCPUtime, 1, 0);
            writime("MD", istep, 1, 0);
            statime(0);
            info_thermo(0, NULL);
        }
        if (iopnode(istep, npas, fprint)) info_thermo(1, fpOSZIFF);
        if (iopnode(istep, npas, cprint)) {
            state_down(c);
            config_out("CONTFF", "w");
        }
        if (tcap > 0 && tpend == tcap) {
            traj_flush(c);
            tpend = 0;
        }
        first = last + 1;
    }
    if (tcap > 0) {
        traj_flush(c);
        if ((rc = mdff_set_trajectory(c, 0, 0, 0, 0))) fail("mdff_set_trajectory", rc);
    }
    state_down(c);
    config_out("CONTFF", "w");
    fclose(fpOSZIFF);
    /* the SCF's dipole Ewald calls of all steps: one static-field call per
     * force call, one per iteration and per inner scfkO sweep */
    long long k1[4];
    if ((rc = mdff_get_counters(c, k1))) fail("mdff_get_counters", rc);
    if (lpim) {
        const long long calls = (k1[0] - k0[0]) + (k1[1] - k0[1]) + (k1[2] - k0[2]);
        cc_multipole_ES_dir_pim += (int)calls;
        cc_multipole_ES_rec_pim += (int)calls;
    }
}

static void multipole_ES_impl(double *q, double (*mu)[3], double (*theta)[3][3], double *u, double *pvir,
                              double tau[3][3], double (*ef)[3], double (*efg)[3][3], bool lq, bool ld, bool lt,
                              bool ldamp, bool do_forces, bool do_stress, bool do_ef, bool do_efg, bool do_dir,
                              bool do_rec, bool inpim, bool update_sf)
{
    mdff_ctx *c = ctx_sync();
    int flags = (lq ? MDFF_ES_LQ : 0) | (ld ? MDFF_ES_LD : 0) | (lt ? MDFF_ES_LT : 0) | (ldamp ? MDFF_ES_LDAMP : 0) |
                (do_forces ? MDFF_ES_FORCES : 0) | (do_stress ? MDFF_ES_STRESS : 0) | (do_ef ? MDFF_ES_EF : 0) |
                (do_efg ? MDFF_ES_EFG : 0) | (do_dir ? MDFF_ES_DIR : 0) | (do_rec ? MDFF_ES_REC : 0) |
                (inpim ? MDFF_ES_INPIM : 0) | (update_sf ? MDFF_ES_UPDATE_SF : 0);
    double t9[9];
    int rc = mdff_multipole_es(c, q, &mu[0][0], theta ? &theta[0][0][0] : NULL, flags, u, pvir, t9,
                               ef ? &ef[0][0] : NULL, efg ? &efg[0][0][0] : NULL, do_forces ? fx : NULL,
                               do_forces ? fy : NULL, do_forces ? fz : NULL);
    if (rc) fail("mdff_multipole_es", rc);
    for (int i = 0; i < 9; i++) tau[i / 3][i % 3] = t9[i];
}

static void get_dipoles_impl(double (*mu)[3], double *upol)
{
    mdff_ctx *c = ctx_sync();
    int rc = mdff_get_dipoles_scf(c, istep, &mu[0][0], upol);
    if (rc && rc != MDFF_ERR_SCF) fail("mdff_get_dipoles_scf", rc);
    if (lpol) ldip = true;
}

static void momentpolaSCF_impl(double (*mu)[3], double *upol)
{
    mdff_ctx *c = ctx_sync();
    int rc = mdff_get_dipoles_scf(c, istep, &mu[0][0], upol);
    if (rc && rc != MDFF_ERR_SCF) fail("mdff_get_dipoles_scf", rc);
    for (int ia = 0; ia < nion; ia++)
        for (int i = 0; i < 3; i++) mu[ia][i] -= dipia[ia][i];
}

static void pair_part(double *u, double *pvir, double tau[3][3])
{
    mdff_ctx *c = ctx_sync();
    double t9[9];
    int rc = mdff_pair_short(c, u, pvir, t9);
    if (rc) fail("mdff_pair_short", rc);
    for (int i = 0; i < 9; i++) tau[i / 3][i % 3] = t9[i];
    double *f = (double *)malloc(sizeof(double) * 3 * (size_t)nion);
    if (!f) fail("malloc", MDFF_ERR_NOMEM);
    rc = mdff_get_forces(c, f, f + nion, f + 2 * nion);
    if (rc) fail("mdff_get_forces", rc);
    for (int ia = 0; ia < nion; ia++) {
        fx[ia] += f[ia];
        fy[ia] += f[nion + ia];
        fz[ia] += f[2 * nion + ia];
    }
    free(f);
}

void engforce_bhmftd_pbc(double *u, double *pvir, double tau[3][3])
{
    char *rs = rng_enter();
    pair_part(u, pvir, tau);
    rng_leave(rs);
}

void engforce_nmlj_pbc(double *u, double *pvir, double tau[3][3])
{
    char *rs = rng_enter();
    pair_part(u, pvir, tau);
    rng_leave(rs);
}

void TT_damping_functions(double b, double c, double r, double *f, double *fd, int order)
{
    mdff_tt_damping(b, c, r, f, fd, order);
}

static void check_verletlist_impl(void)
{
    mdff_ctx *c = ctx_sync();
    int rc = mdff_check_verlet(c);
    if (rc) fail("mdff_check_verlet", rc);
    mdff_thermo th;
    rc = mdff_get_thermo(c, &th);
    if (rc) fail("mdff_get_thermo", rc);
    updatevl = th.verlet_updates;
}

void gen_pbc_verletlist(void) { check_verletlist(); }

/* invertmatrix3x3 inverse_matrix.c:6-23 (LAPACK dgetrf/dgetri there): the
 * inverse by the adjugate, in place; a singular matrix gives non-finite
 * entries, where LAPACK's result is not reproduced */
void invertmatrix3x3(double A[3][3])
{
    double m[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) m[i][j] = A[i][j];
    const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    const double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    const double id = 1.0 / det;
    A[0][0] = c00 * id;
    A[1][0] = c01 * id;
    A[2][0] = c02 * id;
    A[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    A[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    A[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
    A[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    A[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
    A[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
}

/* ---- the same-name entry points: the bodies above on a private rand() state */
void engforce(void)
{
    char *rs = rng_enter();
    engforce_impl();
    rng_leave(rs);
}

void multipole_ES(double *q, double (*mu)[3], double (*theta)[3][3], double *u, double *pvir, double tau[3][3],
                  double (*ef)[3], double (*efg)[3][3], bool lq, bool ld, bool lt, bool ldamp, bool do_forces,
                  bool do_stress, bool do_ef, bool do_efg, bool do_dir, bool do_rec, bool inpim, bool update_sf)
{
    char *rs = rng_enter();
    multipole_ES_impl(q, mu, theta, u, pvir, tau, ef, efg, lq, ld, lt, ldamp, do_forces, do_stress, do_ef, do_efg,
                      do_dir, do_rec, inpim, update_sf);
    rng_leave(rs);
}

void get_dipoles(double (*mu)[3], double *upol)
{
    char *rs = rng_enter();
    get_dipoles_impl(mu, upol);
    rng_leave(rs);
}

void momentpolaSCF(double (*mu)[3], double *upol)
{
    char *rs = rng_enter();
    momentpolaSCF_impl(mu, upol);
    rng_leave(rs);
}

/* get_gr radial_distribution.c:180-231: one frame of the globals on the
 * device; gr gains the ordered pair counts per (bin, type, type) and
 * gr[.][0][0] also every pair (:219-220); rx/ry/rz come back as kardir and
 * dirkar leave them (:194, :230) */
static void get_gr_impl(void)
{
    int same = g_rdf && g_rdf_n == nion && g_rdf_nt == ntype && g_rdf_nb == nbins && g_rdf_resg == resg &&
               g_rdf_cut == cutgr && memcmp(g_rdf_type, typia, sizeof(int) * nion) == 0;
    if (!same) {
        if (g_rdf) mdff_rdf_destroy(g_rdf);
        g_rdf = NULL;
        free(g_rdf_type);
        g_rdf_type = malloc(sizeof(int) * nion);
        if (!g_rdf_type) fail("get_gr (host memory)", MDFF_ERR_NOMEM);
        memcpy(g_rdf_type, typia, sizeof(int) * nion);
        int st = 0;
        g_rdf = mdff_rdf_create(nion, ntype, typia, nbins, resg, cutgr, &st);
        if (!g_rdf) {
            /* nbins of default_rdf (resg 0.1) does not hold the distances
             * below cutgr at a smaller resg; the reference writes past gr */
            fprintf(stderr, "mdff drop-in: get_gr: %d bins of %g A do not reach cutgr %g A\n", nbins, resg, cutgr);
            fail("mdff_rdf_create", st);
        }
        g_rdf_n = nion;
        g_rdf_nt = ntype;
        g_rdf_nb = nbins;
        g_rdf_resg = resg;
        g_rdf_cut = cutgr;
        if (!g_atexit) {
            atexit(mdff_dropin_release);
            g_atexit = 1;
        }
    }
    int rc = mdff_rdf_reset(g_rdf);
    if (rc) fail("mdff_rdf_reset", rc);
    if ((rc = mdff_rdf_add_frame(g_rdf, simuCell.A, rx, ry, rz, 1))) fail("mdff_rdf_add_frame", rc);
    const size_t nc = (size_t)nbins * ntype * ntype;
    long long *cnt = malloc(sizeof(long long) * nc);
    if (!cnt) fail("get_gr (host memory)", MDFF_ERR_NOMEM);
    if ((rc = mdff_rdf_counts(g_rdf, cnt, NULL))) fail("mdff_rdf_counts", rc);
    for (int b = 0; b < nbins; b++) {
        long long all = 0;
        for (int i = 0; i < ntype; i++)
            for (int j = 0; j < ntype; j++) {
                const long long v = cnt[((size_t)b * ntype + i) * ntype + j];
                gr[b][i][j] += (int)v;
                all += v;
            }
        gr[b][0][0] += (int)all;
    }
    free(cnt);
}

void get_gr(void)
{
    char *rs = rng_enter();
    get_gr_impl();
    rng_leave(rs);
}

void check_verletlist(void)
{
    char *rs = rng_enter();
    check_verletlist_impl();
    rng_leave(rs);
}

void prop_agate(int step)
{
    char *rs = rng_enter();
    prop_agate_impl(step);
    rng_leave(rs);
}

void run_md(void)
{
    char *rs = rng_enter();
    run_md_impl();
    rng_leave(rs);
}
