/*
 * control.c -- control-file parser for the force/integration path.
 *
 * The reference re-scans the control file once per module, matching each
 * whitespace-separated token against that module's keywords and reading the
 * values that follow with fscanf (global.c:23-57, md.c:42-97,
 * field.c:50-148, bhmftd.c:34-88, nmlj.c:47-104, nonbonded.c:17-22,
 * pim.c:40-67).  Values are numbers or booleans, never keywords, so one
 * pass applying every module's handlers reads the same values.  A numeric
 * read that fails (for example a short pair-table row) leaves the target
 * unchanged and consumes nothing, as a failed fscanf("%lf") does.
 */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mdff_host.h"

/* accepted boolean spellings, index parity = truth (constants.c:18-33,
 * check_boolstring tools.c:78-80) */
static const char *kBoolStr[16] = {"yes", "no", "y", "n", "Y", "N", "Yes", "No",
                                   "true", "false", "t", "f", "T", "F", "True", "False"};

typedef struct {
    char **tok;
    int n, pos;
} tokstream;

static int load_tokens(const char *path, tokstream *ts)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return MDFF_ERR_IO;
    int cap = 256;
    ts->tok = (char **)malloc(cap * sizeof(char *));
    ts->n = 0;
    ts->pos = 0;
    char buf[4096];
    while (fscanf(fp, "%4095s", buf) == 1) {
        if (ts->n == cap) {
            cap *= 2;
            ts->tok = (char **)realloc(ts->tok, cap * sizeof(char *));
        }
        ts->tok[ts->n++] = strdup(buf);
    }
    fclose(fp);
    return MDFF_OK;
}

static void free_tokens(tokstream *ts)
{
    for (int i = 0; i < ts->n; i++) free(ts->tok[i]);
    free(ts->tok);
}

/* fscanf("%lf")-like: succeeds only on a token that is a complete number */
static int next_double(tokstream *ts, double *v)
{
    if (ts->pos >= ts->n) return 0;
    char *end;
    const char *s = ts->tok[ts->pos];
    double x = strtod(s, &end);
    if (end == s || *end != '\0') return 0;
    *v = x;
    ts->pos++;
    return 1;
}

static int next_int(tokstream *ts, int *v)
{
    if (ts->pos >= ts->n) return 0;
    char *end;
    const char *s = ts->tok[ts->pos];
    long x = strtol(s, &end, 10);
    if (end == s || *end != '\0') return 0;
    *v = (int)x;
    ts->pos++;
    return 1;
}

static const char *next_str(tokstream *ts)
{
    if (ts->pos >= ts->n) return NULL;
    return ts->tok[ts->pos++];
}

/* check_string tools.c:56-68: index of the match, -1 if not allowed */
static int match(const char *s, const char *const *allowed, int n)
{
    int k = -1;
    for (int i = 0; s && i < n; i++)
        if (strcmp(s, allowed[i]) == 0) k = i;
    return k;
}

static int read_bool(tokstream *ts, int *out, const char *label)
{
    const char *s = next_str(ts);
    int k = match(s, kBoolStr, 16);
    if (k < 0) {
        fprintf(stderr, "mdff: value not allowed for %s: %s\n", label, s ? s : "(eof)");
        return MDFF_ERR_CONFIG;
    }
    *out = (k % 2 == 0);
    return MDFF_OK;
}

/* upper-triangular pair table, row-major (bhmftd.c:39-47) */
static void read_pair_table(tokstream *ts, int ntype, double t[MDFF_NTYPEMAX][MDFF_NTYPEMAX])
{
    for (int it = 0; it < ntype; it++)
        for (int jt = it; jt < ntype; jt++) {
            double v;
            if (next_double(ts, &v)) t[it][jt] = v;
        }
}

void mdff_params_default(mdff_params *p)
{
    memset(p, 0, sizeof(*p));
    /* global.c:83-94 */
    p->Fposff = 2;
    p->skindiff = 0.15;
    p->lverletL = 1;
    /* md.c:100-118 */
    p->integrator = MDFF_NVE_VV;
    p->nprint = 1;
    p->fprint = 1;
    p->cprint = 1000;
    p->tprint = 1000;
    p->tstart = 0;
    /* pim.c:72-85 */
    p->conv_tol_ind = 1e-6;
    p->min_scf_pim_iter = 3;
    p->max_scf_pim_iter = 100;
    p->extrapolate_order = 0;
    p->algo_extrapolate_dipole = 0;
    p->algo_pim = MDFF_PIM_SCF;
    /* nmlj.c:110-120 (applied to every pair before the table is read) */
    for (int i = 0; i < MDFF_NTYPEMAX; i++)
        for (int j = 0; j < MDFF_NTYPEMAX; j++) {
            p->plj[i][j] = 6.0;
            p->qlj[i][j] = 12.0;
            p->sigmalj[i][j] = 3.405;
            p->epslj[i][j] = .010323576;
        }
    /* pre_alloc_config config.c:26-33 */
    for (int i = 0; i < MDFF_NTYPEMAX; i++) p->massit[i] = 1.0;
}

int mdff_read_control(const char *path, int ntype, mdff_params *p)
{
    static const char *const integ[3] = {"nve-lf", "nve-vv", "nvt-nhcn"};
    static const char *const fposff[3] = {"rnn", "rvn", "rvf"};
    static const char *const trunc[3] = {"no", "linear", "quadratic"};
    static const char *const algopim[3] = {"scf", "scfkO", "pcg"};
    static const char *const algoext[2] = {"poly", "aspc"};

    if (ntype < 1 || ntype > MDFF_NTYPEMAX) return MDFF_ERR_ARG;
    mdff_params_default(p);
    p->ntype = ntype;
    tokstream ts;
    int rc = load_tokens(path, &ts);
    if (rc) return rc;

    int kmas = 0, kqch = 0, kdip = 0, kqua = 0, kpol = 0;
    int have_trunc = 0;
    while (rc == MDFF_OK && ts.pos < ts.n) {
        const char *b = ts.tok[ts.pos++];
        int k;
        /* ---------------- global.c:24-57 ---------------- */
        if (!strcmp(b, "lverletL")) rc = read_bool(&ts, &p->lverletL, b);
        else if (!strcmp(b, "lstatic")) rc = read_bool(&ts, &p->lstatic, b);
        else if (!strcmp(b, "lrdf")) rc = read_bool(&ts, &p->lrdf, b);
        else if (!strcmp(b, "lpstress")) rc = read_bool(&ts, &p->lpstress, b);
        else if (!strcmp(b, "cutshortrange")) next_double(&ts, &p->cutshortrange);
        else if (!strcmp(b, "cutlongrange")) next_double(&ts, &p->cutlongrange);
        else if (!strcmp(b, "Fposff")) {
            k = match(next_str(&ts), fposff, 3);
            if (k < 0) rc = MDFF_ERR_CONFIG; else p->Fposff = k;
        } else if (!strcmp(b, "skindiff")) next_double(&ts, &p->skindiff);
        else if (!strcmp(b, "lreduced")) rc = read_bool(&ts, &p->lreduced, b);
        /* ---------------- md.c:42-97 ---------------- */
        else if (!strcmp(b, "integrator")) {
            k = match(next_str(&ts), integ, 3);
            if (k < 0) rc = MDFF_ERR_CONFIG; else p->integrator = k;
        } else if (!strcmp(b, "dt")) next_double(&ts, &p->dt);
        else if (!strcmp(b, "temp")) next_double(&ts, &p->temp);
        else if (!strcmp(b, "npas")) next_int(&ts, &p->npas);
        else if (!strcmp(b, "nprint")) next_int(&ts, &p->nprint);
        else if (!strcmp(b, "fprint")) next_int(&ts, &p->fprint);
        else if (!strcmp(b, "cprint")) next_int(&ts, &p->cprint);
        else if (!strcmp(b, "tstart")) next_int(&ts, &p->tstart);
        else if (!strcmp(b, "tprint")) next_int(&ts, &p->tprint);
        else if (!strcmp(b, "nequil")) next_int(&ts, &p->nequil);
        else if (!strcmp(b, "nequilT")) next_int(&ts, &p->nequilT);
        else if (!strcmp(b, "tauTberendsen")) next_double(&ts, &p->tauTberendsen);
        else if (!strcmp(b, "nhc_n")) next_int(&ts, &p->nhc_n);
        else if (!strcmp(b, "nhc_mults")) next_int(&ts, &p->nhc_mults);
        else if (!strcmp(b, "nhc_yosh_order")) next_int(&ts, &p->nhc_yosh_order);
        else if (!strcmp(b, "timesca_thermo")) next_double(&ts, &p->timesca_thermo);
        else if (!strcmp(b, "timesca_baro")) next_double(&ts, &p->timesca_baro);
        /* ---------------- field.c:50-148 ---------------- */
        else if (!strcmp(b, "lbhmft")) rc = read_bool(&ts, &p->lbhmft, b);
        else if (!strcmp(b, "lbhmftd")) rc = read_bool(&ts, &p->lbhmftd, b);
        else if (!strcmp(b, "lnmlj")) rc = read_bool(&ts, &p->lnmlj, b);
        else if (!strcmp(b, "lcoulombic")) rc = read_bool(&ts, &p->lcoulombic, b);
        else if (!strcmp(b, "massit")) {
            if (kmas < MDFF_NTYPEMAX) next_double(&ts, &p->massit[kmas]);
            kmas++;
        } else if (!strcmp(b, "qit")) {
            if (kqch < MDFF_NTYPEMAX) next_double(&ts, &p->qit[kqch]);
            kqch++;
        } else if (!strcmp(b, "dipit")) {
            for (int i = 0; i < 3; i++)
                if (kdip < MDFF_NTYPEMAX) next_double(&ts, &p->dipit[kdip][i]);
            kdip++;
        } else if (!strcmp(b, "quadit")) {
            for (int i = 0; i < 9; i++)
                if (kqua < MDFF_NTYPEMAX) next_double(&ts, &p->quadit[kqua][i / 3][i % 3]);
            kqua++;
        } else if (!strcmp(b, "polit")) {
            if (kpol < MDFF_NTYPEMAX) {
                for (int i = 0; i < 9; i++) next_double(&ts, &p->polit[kpol][i / 3][i % 3]);
                for (int i = 0; i < 9; i++)
                    if (p->polit[kpol][i / 3][i % 3] != 0.0) p->lpolar[kpol] = 1;
            }
            kpol++;
        } else if (!strcmp(b, "lpoldamping")) {
            int t = 0, it = 0, jt = 0, v = 0;
            next_int(&ts, &t); next_int(&ts, &it); next_int(&ts, &jt);
            rc = read_bool(&ts, &v, b);
            if (t < 0 || t >= MDFF_NTYPEMAX || it < 0 || it >= MDFF_NTYPEMAX || jt < 0 || jt >= MDFF_NTYPEMAX)
                rc = MDFF_ERR_CONFIG;
            else p->lpoldamping[t][it][jt] = v;
        } else if (!strcmp(b, "pol_damp_b") || !strcmp(b, "pol_damp_c") || !strcmp(b, "pol_damp_k")) {
            int t = 0, it = 0, jt = 0;
            next_int(&ts, &t); next_int(&ts, &it); next_int(&ts, &jt);
            if (t < 0 || t >= MDFF_NTYPEMAX || it < 0 || it >= MDFF_NTYPEMAX || jt < 0 || jt >= MDFF_NTYPEMAX) {
                rc = MDFF_ERR_CONFIG;
            } else if (b[9] == 'b') next_double(&ts, &p->pol_damp_b[t][it][jt]);
            else if (b[9] == 'c') next_double(&ts, &p->pol_damp_c[t][it][jt]);
            else next_int(&ts, &p->pol_damp_k[t][it][jt]);
        } else if (!strcmp(b, "alphaES")) next_double(&ts, &p->alphaES);
        else if (!strcmp(b, "kES")) {
            for (int i = 0; i < 3; i++) next_int(&ts, &p->kES[i]);
        } else if (!strcmp(b, "epsw")) {
            /* read_field tests "epsw" twice (field.c:129-134): a second
             * number directly after the first overwrites it */
            next_double(&ts, &p->epsw);
            next_double(&ts, &p->epsw);
        } else if (!strcmp(b, "lautoES")) rc = read_bool(&ts, &p->lautoES, b);
        /* ---------------- bhmftd.c:34-88 ---------------- */
        else if (!strcmp(b, "Abhmftd")) read_pair_table(&ts, ntype, p->Abhmftd);
        else if (!strcmp(b, "Bbhmftd")) read_pair_table(&ts, ntype, p->Bbhmftd);
        else if (!strcmp(b, "Cbhmftd")) read_pair_table(&ts, ntype, p->Cbhmftd);
        else if (!strcmp(b, "Dbhmftd")) read_pair_table(&ts, ntype, p->Dbhmftd);
        else if (!strcmp(b, "BDbhmftd")) read_pair_table(&ts, ntype, p->BDbhmftd);
        /* ---------------- nmlj.c:47-104 ---------------- */
        else if (!strcmp(b, "sigmalj")) read_pair_table(&ts, ntype, p->sigmalj);
        else if (!strcmp(b, "epslj")) read_pair_table(&ts, ntype, p->epslj);
        else if (!strcmp(b, "plj")) read_pair_table(&ts, ntype, p->plj);
        else if (!strcmp(b, "qlj")) read_pair_table(&ts, ntype, p->qlj);
        else if (!strcmp(b, "trunc")) {
            k = match(next_str(&ts), trunc, 3);
            if (k < 0) rc = MDFF_ERR_CONFIG; else { p->trunc = k; have_trunc = 1; }
        }
        /* ---------------- nonbonded.c:17-22 ---------------- */
        else if (!strcmp(b, "lsymmetric")) p->lsymmetric = 1;
        /* ---------------- pim.c:40-67 ---------------- */
        else if (!strcmp(b, "algo_pim")) {
            k = match(next_str(&ts), algopim, 3);
            if (k < 0) rc = MDFF_ERR_CONFIG; else p->algo_pim = k;
        } else if (!strcmp(b, "algo_extrapolate_dipole")) {
            k = match(next_str(&ts), algoext, 2);
            if (k < 0) rc = MDFF_ERR_CONFIG; else p->algo_extrapolate_dipole = k;
        } else if (!strcmp(b, "conv_tol_ind")) next_double(&ts, &p->conv_tol_ind);
        else if (!strcmp(b, "extrapolate_order")) next_int(&ts, &p->extrapolate_order);
        else if (!strcmp(b, "min_scf_pim_iter")) next_int(&ts, &p->min_scf_pim_iter);
        else if (!strcmp(b, "max_scf_pim_iter")) next_int(&ts, &p->max_scf_pim_iter);
        else if (!strcmp(b, "omegakO")) next_double(&ts, &p->omegakO);
    }
    free_tokens(&ts);
    (void)have_trunc;
    if (rc) return rc;
    p->nmassit = kmas; p->nqit = kqch; p->ndipit = kdip; p->nquadit = kqua; p->npolit = kpol;
    /* field.c:150-154 */
    if (kmas != ntype && kmas != 0) return MDFF_ERR_CONFIG;

    /* symmetric pair tables (bhmftd.c:98-110, nmlj.c:205-212): the upper
     * triangle read from the file is copied to the lower one */
    for (int it = 0; it < ntype; it++)
        for (int jt = it + 1; jt < ntype; jt++) {
            p->Abhmftd[jt][it] = p->Abhmftd[it][jt];
            p->Bbhmftd[jt][it] = p->Bbhmftd[it][jt];
            p->Cbhmftd[jt][it] = p->Cbhmftd[it][jt];
            p->Dbhmftd[jt][it] = p->Dbhmftd[it][jt];
            p->BDbhmftd[jt][it] = p->BDbhmftd[it][jt];
            p->plj[jt][it] = p->plj[it][jt];
            p->qlj[jt][it] = p->qlj[it][jt];
            p->sigmalj[jt][it] = p->sigmalj[it][jt];
            p->epslj[jt][it] = p->epslj[it][jt];
        }
    /* damping tables: field.c:209-219 copies [t][it][kt] (it<kt) onto
     * [t][kt][it] in loop order, so the upper entry wins */
    for (int t = 0; t < ntype; t++)
        for (int it = 0; it < ntype; it++)
            for (int kt = it + 1; kt < ntype; kt++) {
                p->lpoldamping[t][kt][it] = p->lpoldamping[t][it][kt];
                p->pol_damp_b[t][kt][it] = p->pol_damp_b[t][it][kt];
                p->pol_damp_c[t][kt][it] = p->pol_damp_c[t][it][kt];
                p->pol_damp_k[t][kt][it] = p->pol_damp_k[t][it][kt];
            }
    /* pol_damp_k indexes E_TT[0..8] (tt_damp.h:4) */
    for (int t = 0; t < ntype; t++)
        for (int it = 0; it < ntype; it++)
            for (int kt = 0; kt < ntype; kt++)
                if (p->pol_damp_k[t][it][kt] < 0 || p->pol_damp_k[t][it][kt] > MDFF_MAX_TT_ORDER)
                    return MDFF_ERR_CONFIG;
    if (p->lreduced) { /* nothing to convert: units are applied by the engines */ }
    return MDFF_OK;
}

void mdff_units_get(const mdff_params *p, mdff_units *u)
{
    if (p && p->lreduced) {
        u->boltz_unit = u->time_unit = u->press_unit = u->rho_unit = u->coul_unit = 1.0;
        return;
    }
    u->boltz_unit = 8.61734229648141E-05;
    u->time_unit = 98.2269514139276;
    u->press_unit = 0.00624150964712042;
    u->rho_unit = 1.660538782;
    u->coul_unit = 14.3996441494161;
}

size_t mdff_sizeof_params(void) { return sizeof(mdff_params); }
size_t mdff_sizeof_thermo(void) { return sizeof(mdff_thermo); }
size_t mdff_sizeof_config(void) { return sizeof(mdff_config); }
