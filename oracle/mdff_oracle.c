/*
 * mdff_oracle.c -- CPU FP64 restatement of mdff20's force and integration
 * path (TEST INFRASTRUCTURE: the checker for the HIP engine, never part of
 * the product path).
 *
 * Pinned by: the golden vectors written by the reference build
 * (tests/golden/, made by tests/golden/make_fixtures.py through
 * oracle/_ref), the CCD Lennard-Jones cluster energies of the reference's
 * ex2 example, Madelung constants, closed forms for two charges / one
 * dipole / one polarisable ion, and finite differences (tests/test_oracle*.py).
 *
 * Structure follows the reference routine by routine; each function cites
 * the reference file:line it restates.  Deliberate differences:
 *  - positions are not round-tripped cart->frac->cart inside every force
 *    routine (Q15, bhmftd.c:139,266): frac copies are made instead, which
 *    changes positions by no more than an ulp per call in the reference;
 *  - summation order of OpenMP reductions differs (the reference's own
 *    order is not reproducible across runs, schedule(dynamic));
 *  - ASPC predictor coefficients are fixed per call (see aspc_predict).
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "mdff_oracle.h"

#define PI 3.1415926535897932385
#define TPI 6.2831853071795864770
#define NT MDFF_NTYPEMAX

typedef struct {
    double boltz, time, press, rho, coul;
} units_t;

typedef struct {
    double A[3][3], B[3][3], Omega, inveOmega, inveOmegaPU, Anorm[3], w[3];
} cell_t;

typedef struct {
    int *point, *list;
    int cap;
    double cut;
} vlist_t;

struct orc_sys {
    mdff_params p;
    units_t u;
    cell_t cell;
    int nion, ntype;
    int *typ;
    double *rx, *ry, *rz, *vx, *vy, *vz, *fx, *fy, *fz, *rxs, *rys, *rzs;
    double *mass, *invmass, *q;
    double (*dip)[3];
    double (*pol)[3][3], (*invpol)[3][3];
    double (*hist)[MDFF_MAX_EXTRAPOLATE_ORDER + 1][3];
    double (*mu)[3];
    /* derived field settings (field.c:186-335) */
    int lqch, ldip, lqua, lpol, ldmp, lpim, lnonbonded, integ, rescale_allowed;
    double srcutsq, lrcutsq, alpha;
    int kES[3];
    /* k mesh (kspace.c:14-102) */
    int nk;
    double *kx, *ky, *kz, *kk, *Ak, *kcoe;
    double *ckr, *skr;      /* cos/sin(k.r) [nk][N], valid for pos_version */
    long pos_version, table_version;
    /* Verlet lists (verlet.h:8-17) */
    vlist_t vnb, vcoul;
    double *xvl, *yvl, *zvl;
    int updatevl, have_list;
    /* MD */
    double kT, dt, tauT, timesca;
    double *vxi, *xi;
    double E_TT[MDFF_MAX_TT_ORDER + 1];
    mdff_thermo th;
    int istep;
    int ko_inner;   /* scfkO: inner sweeps of the last SCF */
};

/* ---------------------------------------------------------------- units */
/* gen_constants constants.c:8-36, reduced_units constants.c:38-45 */
static void units_init(units_t *u, int lreduced)
{
    if (lreduced) {
        u->boltz = u->time = u->press = u->rho = u->coul = 1.0;
        return;
    }
    u->boltz = 8.61734229648141E-05;
    u->time = 98.2269514139276;
    u->press = 0.00624150964712042;
    u->rho = 1.660538782;
    u->coul = 14.3996441494161;
}

/* ----------------------------------------------------------------- cell */
static void cross3(double h[3], const double *a, const double *b)
{
    h[0] = a[1] * b[2] - a[2] * b[1];
    h[1] = a[2] * b[0] - a[0] * b[2];
    h[2] = a[0] * b[1] - a[1] * b[0];
}

/* lattice cell.c:13-91 (volume formula cell.c:38-41) */
static void cell_init(cell_t *c, const double *A9, double press)
{
    memset(c, 0, sizeof(*c));
    for (int i = 0; i < 9; i++) c->A[i / 3][i % 3] = A9[i];
    cross3(c->B[0], c->A[1], c->A[2]);
    cross3(c->B[1], c->A[2], c->A[0]);
    cross3(c->B[2], c->A[0], c->A[1]);
    double om = c->B[0][0] * c->A[0][0] + c->B[1][0] * c->A[1][0] + c->B[2][0] * c->A[2][0];
    c->Omega = om;
    c->inveOmega = 1.0 / om;
    c->inveOmegaPU = 1.0 / (om * press);
    for (int i = 0; i < 3; i++)
        c->w[i] = om / sqrt(c->B[0][i] * c->B[0][i] + c->B[1][i] * c->B[1][i] + c->B[2][i] * c->B[2][i]);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) c->B[i][j] *= c->inveOmega;
    for (int i = 0; i < 3; i++) {
        double a = 0;
        for (int j = 0; j < 3; j++) a += c->A[j][i] * c->A[j][i];
        c->Anorm[i] = sqrt(a);
    }
}

/* kardir cell.c:125-143 */
static inline void frac(const cell_t *c, double x, double y, double z, double *s)
{
    s[0] = x * c->B[0][0] + y * c->B[0][1] + z * c->B[0][2];
    s[1] = x * c->B[1][0] + y * c->B[1][1] + z * c->B[1][2];
    s[2] = x * c->B[2][0] + y * c->B[2][1] + z * c->B[2][2];
}

/* pbc pbc.c:7-14 on a fractional difference -> cartesian minimum image */
static inline void min_image(const cell_t *c, double sx, double sy, double sz, double *r)
{
    sx -= nearbyint(sx);
    sy -= nearbyint(sy);
    sz -= nearbyint(sz);
    r[0] = sx * c->A[0][0] + sy * c->A[1][0] + sz * c->A[2][0];
    r[1] = sx * c->A[0][1] + sy * c->A[1][1] + sz * c->A[2][1];
    r[2] = sx * c->A[0][2] + sy * c->A[1][2] + sz * c->A[2][2];
}

/* ------------------------------------------------------ Tang-Toennies */
/* get_TT_damp tt_damp.c:34-44 */
static void tt_init(double *E)
{
    E[0] = 1.0;
    for (int i = 1; i <= MDFF_MAX_TT_ORDER; i++) E[i] = E[i - 1] / (double)i;
}

/* TT_damping_functions tt_damp.c:16-32 */
static inline void tt_damp(const double *E, double b, double c, double r, double *f, double *fd, int order)
{
    double br = b * r;
    double ex = exp(-br) * c;
    double ff = E[order];
    for (int k = order - 1; k > -1; k--) ff = ff * br + E[k];
    *f = 1.0 - ff * ex;
    *fd = E[order] * pow(br, order) * ex * b;
}

/* --------------------------------------------------------------- k mesh */
static int cmp_kk(const void *a, const void *b)
{
    const double *x = (const double *)a, *y = (const double *)b;
    return (x[0] < y[0]) ? -1 : (x[0] > y[0]) ? 1 : 0;
}

/* init_kspace kspace.c:14-41, set_param_kmesh :64-102, reorder_kmesh
 * :106-150 (sorted by |k|^2; ties keep enumeration order here) */
static void kmesh_init(orc_sys *s)
{
    int K0 = s->kES[0], K1 = s->kES[1], K2 = s->kES[2];
    int nk = K2 + K1 * (2 * K2 + 1) + K0 * (2 * K1 + 1) * (2 * K2 + 1);
    double *tmp = (double *)malloc(sizeof(double) * 7 * (nk > 0 ? nk : 1));
    double a2 = s->alpha * s->alpha;
    int n = 0;
    for (int nx = 0; nx <= K0; nx++) {
        int nymin = (nx == 0) ? 0 : -K1;
        for (int ny = nymin; ny <= K1; ny++) {
            int nzmin = (nx == 0 && ny == 0) ? 1 : -K2;
            for (int nz = nzmin; nz <= K2; nz++) {
                const double (*B)[3] = (const double(*)[3])s->cell.B;
                double kx = TPI * ((double)nx * B[0][0] + (double)ny * B[0][1] + (double)nz * B[0][2]);
                double ky = TPI * ((double)nx * B[1][0] + (double)ny * B[1][1] + (double)nz * B[1][2]);
                double kz = TPI * ((double)nx * B[2][0] + (double)ny * B[2][1] + (double)nz * B[2][2]);
                double kk = kx * kx + ky * ky + kz * kz;
                double *t = tmp + 7 * n;
                t[0] = kk; t[1] = kx; t[2] = ky; t[3] = kz;
                t[4] = exp(-kk * 0.25 / a2) / kk;
                t[5] = 2.0 * (1.0 / kk + 1.0 / a2 / 4.0);
                t[6] = (double)n;
                n++;
            }
        }
    }
    qsort(tmp, n, 7 * sizeof(double), cmp_kk);
    s->nk = n;
    s->kx = (double *)malloc(sizeof(double) * (n + 1));
    s->ky = (double *)malloc(sizeof(double) * (n + 1));
    s->kz = (double *)malloc(sizeof(double) * (n + 1));
    s->kk = (double *)malloc(sizeof(double) * (n + 1));
    s->Ak = (double *)malloc(sizeof(double) * (n + 1));
    s->kcoe = (double *)malloc(sizeof(double) * (n + 1));
    for (int i = 0; i < n; i++) {
        double *t = tmp + 7 * i;
        s->kk[i] = t[0]; s->kx[i] = t[1]; s->ky[i] = t[2]; s->kz[i] = t[3];
        s->Ak[i] = t[4]; s->kcoe[i] = t[5];
    }
    free(tmp);
}

/* struct_fact kspace.c:154-187: cos/sin(k.r) tables at the current
 * positions (built once per position update; the reference rebuilds them
 * in the first Ewald call of each SCF, pim.c:216-218) */
static void struct_fact(orc_sys *s)
{
    if (s->table_version == s->pos_version && s->ckr) return;
    size_t sz = (size_t)s->nk * (size_t)s->nion;
    if (!s->ckr) {
        s->ckr = (double *)malloc(sizeof(double) * (sz ? sz : 1));
        s->skr = (double *)malloc(sizeof(double) * (sz ? sz : 1));
    }
#pragma omp parallel for schedule(static)
    for (int ik = 0; ik < s->nk; ik++) {
        double kx = s->kx[ik], ky = s->ky[ik], kz = s->kz[ik];
        for (int ia = 0; ia < s->nion; ia++) {
            double kr = kx * s->rx[ia] + ky * s->ry[ia] + kz * s->rz[ia];
            s->ckr[(size_t)ik * s->nion + ia] = cos(kr);
            s->skr[(size_t)ik * s->nion + ia] = sin(kr);
        }
    }
    s->table_version = s->pos_version;
}

/* ----------------------------------------------------------- Verlet list */
static void vlist_push(vlist_t *v, int *count, int ja)
{
    if (*count >= v->cap) {
        v->cap = v->cap * 2 + 1024;
        v->list = (int *)realloc(v->list, sizeof(int) * v->cap);
    }
    v->list[(*count)++] = ja;
}

/* gen_pbc_verletlist verlet.c:54-122: half lists with the parity rule
 * (i>j, i+j even) or (i<j, i+j odd), d^2 <= (rc+skin)^2 */
static void gen_verlet(orc_sys *s)
{
    int n = s->nion;
    double *fr = (double *)malloc(sizeof(double) * 3 * n);
    for (int i = 0; i < n; i++) frac(&s->cell, s->rx[i], s->ry[i], s->rz[i], fr + 3 * i);
    double cnb = s->vnb.cut * s->vnb.cut, cco = s->vcoul.cut * s->vcoul.cut;
    int inb = 0, ico = 0;
    for (int ia = 0; ia < n; ia++) {
        s->vnb.point[ia] = inb;
        s->vcoul.point[ia] = ico;
        for (int ja = 0; ja < n; ja++) {
            if (((ia > ja) && ((ia + ja) % 2 == 0)) || ((ia < ja) && ((ia + ja) % 2 != 0))) {
                double r[3];
                min_image(&s->cell, fr[3 * ia] - fr[3 * ja], fr[3 * ia + 1] - fr[3 * ja + 1],
                          fr[3 * ia + 2] - fr[3 * ja + 2], r);
                double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                if (d2 <= cnb) vlist_push(&s->vnb, &inb, ja);
                if (d2 <= cco) vlist_push(&s->vcoul, &ico, ja);
            }
        }
    }
    s->vnb.point[n] = inb;
    s->vcoul.point[n] = ico;
    free(fr);
}

/* check_verletlist verlet.c:124-170: rebuild when the two largest
 * displacements since the last build exceed the skin.  The first call
 * always builds (the reference compares with uninitialised xvl). */
void orc_check_verletlist(orc_sys *s)
{
    int n = s->nion;
    double m1 = 0.0, m2 = 0.0;
    if (s->have_list) {
        for (int ia = 0; ia < n; ia++) {
            double f[3], r[3];
            frac(&s->cell, s->rx[ia], s->ry[ia], s->rz[ia], f);
            min_image(&s->cell, f[0] - s->xvl[ia], f[1] - s->yvl[ia], f[2] - s->zvl[ia], r);
            double d = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            if (d > m1) { m2 = m1; m1 = d; }
            else if (d > m2) m2 = d;
        }
    }
    if (!s->have_list || m1 + m2 > s->p.skindiff) {
        s->updatevl++;
        gen_verlet(s);
        for (int ia = 0; ia < n; ia++) {
            double f[3];
            frac(&s->cell, s->rx[ia], s->ry[ia], s->rz[ia], f);
            s->xvl[ia] = f[0]; s->yvl[ia] = f[1]; s->zvl[ia] = f[2];
        }
        s->have_list = 1;
    }
    s->th.verlet_updates = s->updatevl;
}

/* -------------------------------------------------------- pair iteration */
/* each unordered pair once: from the half list when lverletL, else j>i
 * (bhmftd.c:156-172, ewald.c:440-465) */
static inline void pair_range(const orc_sys *s, const vlist_t *v, int ia, int *jb, int *je)
{
    if (s->p.lverletL) { *jb = v->point[ia]; *je = v->point[ia + 1]; }
    else { *jb = ia + 1; *je = s->nion; }
}

static inline int pair_j(const orc_sys *s, const vlist_t *v, int j1)
{
    return s->p.lverletL ? v->list[j1] : j1;
}

static int nthreads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static int thread_id(void)
{
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

/* engforce_bhmftd_pbc bhmftd.c:117-271 */
static void engforce_bhmftd(orc_sys *s, double *u, double *pvir, double tau[3][3])
{
    int n = s->nion, nt = nthreads();
    double *F = (double *)calloc((size_t)nt * 3 * n, sizeof(double));
    double *acc = (double *)calloc((size_t)nt * 8, sizeof(double));
    double *fr = (double *)malloc(sizeof(double) * 3 * n);
    for (int i = 0; i < n; i++) frac(&s->cell, s->rx[i], s->ry[i], s->rz[i], fr + 3 * i);
    const mdff_params *p = &s->p;
#pragma omp parallel
    {
        int t = thread_id();
        double *Ft = F + (size_t)t * 3 * n, *at = acc + 8 * t;
#pragma omp for schedule(dynamic, 16)
        for (int ia = 0; ia < n; ia++) {
            int jb, je;
            pair_range(s, &s->vnb, ia, &jb, &je);
            for (int j1 = jb; j1 < je; j1++) {
                int ja = pair_j(s, &s->vnb, j1);
                if (ja == ia) continue;
                int p1 = s->typ[ia], p2 = s->typ[ja];
                if (p->Abhmftd[p1][p2] == 0.0) continue;
                double r[3];
                min_image(&s->cell, fr[3 * ia] - fr[3 * ja], fr[3 * ia + 1] - fr[3 * ja + 1],
                          fr[3 * ia + 2] - fr[3 * ja + 2], r);
                double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                if (!(d2 < s->srcutsq)) continue;
                double ir2 = 1.0 / d2, d = sqrt(d2);
                double erh = p->Abhmftd[p1][p2] * exp(-p->Bbhmftd[p1][p2] * d);
                double ir6 = ir2 * ir2 * ir2;
                double ir8 = ir6 * ir2;
                ir6 *= p->Cbhmftd[p1][p2];
                ir8 *= p->Dbhmftd[p1][p2];
                double f6 = 1.0, f8 = 1.0, fd6 = 0.0, fd8 = 0.0;
                if (p->lbhmftd) {
                    tt_damp(s->E_TT, p->BDbhmftd[p1][p2], 1.0, d, &f6, &fd6, 6);
                    tt_damp(s->E_TT, p->BDbhmftd[p1][p2], 1.0, d, &f8, &fd8, 8);
                }
                double ir6d = ir6 * f6, ir8d = ir8 * f8;
                double ir7 = 6.0 * ir6d / d, ir9 = 8.0 * ir8d / d;
                at[0] += erh - ir6d - ir8d;
                double w = (p->Bbhmftd[p1][p2] * erh - ir7 - ir9 + ir6 * fd6 + ir8 * fd8) / d;
                double fx = w * r[0], fy = w * r[1], fz = w * r[2];
                Ft[3 * ia] += fx; Ft[3 * ia + 1] += fy; Ft[3 * ia + 2] += fz;
                Ft[3 * ja] -= fx; Ft[3 * ja + 1] -= fy; Ft[3 * ja + 2] -= fz;
                at[1] += r[0] * fx; at[2] += (r[0] * fy + r[1] * fx) * 0.5;
                at[3] += (r[0] * fz + r[2] * fx) * 0.5; at[4] += r[1] * fy;
                at[5] += (r[1] * fz + r[2] * fy) * 0.5; at[6] += r[2] * fz;
            }
        }
    }
    double uu = 0, tt[6] = {0};
    for (int t = 0; t < nt; t++) {
        uu += acc[8 * t];
        for (int k = 0; k < 6; k++) tt[k] += acc[8 * t + 1 + k];
        for (int i = 0; i < n; i++) {
            s->fx[i] += F[(size_t)t * 3 * n + 3 * i];
            s->fy[i] += F[(size_t)t * 3 * n + 3 * i + 1];
            s->fz[i] += F[(size_t)t * 3 * n + 3 * i + 2];
        }
    }
    double full[3][3] = {{tt[0], tt[1], tt[2]}, {tt[1], tt[3], tt[4]}, {tt[2], tt[4], tt[5]}};
    *pvir = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) tau[i][j] = full[i][j] * s->cell.inveOmegaPU;
    for (int i = 0; i < 3; i++) *pvir += tau[i][i] * (1.0 / 3.0);
    *u = uu;
    free(F); free(acc); free(fr);
}

/* init_nmlj nmlj.c:123-251 derived tables + engforce_nmlj_pbc nmlj.c:254-385 */
static void engforce_nmlj(orc_sys *s, double *u, double *pvir, double tau[3][3])
{
    const mdff_params *p = &s->p;
    double sigsq[NT][NT], epsqp[NT][NT], fc[NT][NT], ptwo[NT][NT], qtwo[NT][NT], uc[NT][NT];
    double two16 = pow(2.0, 1.0 / 6.0);
    for (int it = 0; it < s->ntype; it++)
        for (int jt = 0; jt < s->ntype; jt++) {
            double pp = p->plj[it][jt], qq = p->qlj[it][jt];
            epsqp[it][jt] = p->epslj[it][jt] / (qq - pp);
            sigsq[it][jt] = two16 * two16 * p->sigmalj[it][jt] * p->sigmalj[it][jt];
            fc[it][jt] = (pp * qq * epsqp[it][jt]) / sigsq[it][jt];
            ptwo[it][jt] = pp * 0.5;
            qtwo[it][jt] = qq * 0.5;
            double sr2 = sigsq[it][jt] / s->srcutsq, sr = sqrt(sr2);
            uc[it][jt] = epsqp[it][jt] * (pp * pow(sr, qq) - qq * pow(sr, pp));
        }
    int n = s->nion, nt = nthreads();
    double *F = (double *)calloc((size_t)nt * 3 * n, sizeof(double));
    double *acc = (double *)calloc((size_t)nt * 8, sizeof(double));
    double *fr = (double *)malloc(sizeof(double) * 3 * n);
    for (int i = 0; i < n; i++) frac(&s->cell, s->rx[i], s->ry[i], s->rz[i], fr + 3 * i);
#pragma omp parallel
    {
        int t = thread_id();
        double *Ft = F + (size_t)t * 3 * n, *at = acc + 8 * t;
#pragma omp for schedule(dynamic, 16)
        for (int ia = 0; ia < n; ia++) {
            int jb, je;
            pair_range(s, &s->vnb, ia, &jb, &je);
            for (int j1 = jb; j1 < je; j1++) {
                int ja = pair_j(s, &s->vnb, j1);
                if (ja == ia) continue;
                double r[3];
                min_image(&s->cell, fr[3 * ia] - fr[3 * ja], fr[3 * ia + 1] - fr[3 * ja + 1],
                          fr[3 * ia + 2] - fr[3 * ja + 2], r);
                double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                if (!(d2 < s->srcutsq)) continue;
                int p1 = s->typ[ia], p2 = s->typ[ja];
                double sr2 = sigsq[p1][p2] / d2;
                double srp = pow(sr2, ptwo[p1][p2]);
                double srq = pow(sr2, qtwo[p1][p2]);
                /* addtruncU nmlj.c:28-36 */
                if (p->trunc == MDFF_TRUNC_NO)
                    at[0] += epsqp[p1][p2] * (p->plj[p1][p2] * srq - p->qlj[p1][p2] * srp);
                else if (p->trunc == MDFF_TRUNC_LINEAR)
                    at[0] += epsqp[p1][p2] * (p->plj[p1][p2] * srq - p->qlj[p1][p2] * srp) - uc[p1][p2];
                double w = fc[p1][p2] * (srq - srp) * sr2;
                double fx = w * r[0], fy = w * r[1], fz = w * r[2];
                Ft[3 * ia] += fx; Ft[3 * ia + 1] += fy; Ft[3 * ia + 2] += fz;
                Ft[3 * ja] -= fx; Ft[3 * ja + 1] -= fy; Ft[3 * ja + 2] -= fz;
                at[1] += r[0] * fx; at[2] += (r[0] * fy + r[1] * fx) * 0.5;
                at[3] += (r[0] * fz + r[2] * fx) * 0.5; at[4] += r[1] * fy;
                at[5] += (r[1] * fz + r[2] * fy) * 0.5; at[6] += r[2] * fz;
            }
        }
    }
    double uu = 0, tt[6] = {0};
    for (int t = 0; t < nt; t++) {
        uu += acc[8 * t];
        for (int k = 0; k < 6; k++) tt[k] += acc[8 * t + 1 + k];
        for (int i = 0; i < n; i++) {
            s->fx[i] += F[(size_t)t * 3 * n + 3 * i];
            s->fy[i] += F[(size_t)t * 3 * n + 3 * i + 1];
            s->fz[i] += F[(size_t)t * 3 * n + 3 * i + 2];
        }
    }
    double full[3][3] = {{tt[0], tt[1], tt[2]}, {tt[1], tt[3], tt[4]}, {tt[2], tt[4], tt[5]}};
    *pvir = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) tau[i][j] = full[i][j] * s->cell.inveOmegaPU;
    for (int i = 0; i < 3; i++) *pvir += tau[i][i] * (1.0 / 3.0);
    *u = uu;
    free(F); free(acc); free(fr);
}

/* ----------------------------------------------------- multipole Ewald */
/* per-thread accumulators of multipole_ES_dir */
typedef struct {
    double *f, *ef, *efg;
    double u, tt[6];
} dir_acc;

/* multipole_ES_dir ewald.c:353-871.  T-tensors of rank 1..3 of the
 * screened Coulomb operator (ewald.c:524-617); rank 4 and 5 tensors are
 * built by the reference but never contracted, so they are not formed.
 * Quadrupoles take no part in any term of the reference (lquatask only
 * builds T5), so theta is not an input. */
static void ewald_dir(orc_sys *s, const double *q, const double (*mu)[3], int flags,
                      double *u_dir, double *ef_dir, double *efg_dir, double *f_dir, double tau_dir[3][3])
{
    const int lq = flags & ORC_LQ, ld = flags & ORC_LD, lcd = flags & ORC_LDAMP;
    const int do_f = flags & ORC_FORCES, do_s = flags & ORC_STRESS;
    const int do_ef = flags & ORC_EF, do_efg = flags & ORC_EFG;
    const int lqd = lq && ld;
    const mdff_params *p = &s->p;
    int n = s->nion, nt = nthreads();
    double a = s->alpha, a2 = a * a, a3 = a2 * a, a5 = a3 * a2;
    double piroot = sqrt(PI);
    dir_acc *acc = (dir_acc *)calloc(nt, sizeof(dir_acc));
    for (int t = 0; t < nt; t++) {
        acc[t].f = (double *)calloc((size_t)3 * n, sizeof(double));
        acc[t].ef = (double *)calloc((size_t)3 * n, sizeof(double));
        acc[t].efg = (double *)calloc((size_t)9 * n, sizeof(double));
    }
    double *fr = (double *)malloc(sizeof(double) * 3 * n);
    for (int i = 0; i < n; i++) frac(&s->cell, s->rx[i], s->ry[i], s->rz[i], fr + 3 * i);
#pragma omp parallel
    {
        dir_acc *A = acc + thread_id();
#pragma omp for schedule(dynamic, 8)
        for (int ia = 0; ia < n; ia++) {
            double qi = q[ia];
            int qch_i = (qi != 0.0);
            const double *mui = mu[ia];
            int dip_i = (mui[0] != 0.0) || (mui[1] != 0.0) || (mui[2] != 0.0);
            int ita = s->typ[ia];
            int jb, je;
            pair_range(s, &s->vcoul, ia, &jb, &je);
            for (int j1 = jb; j1 < je; j1++) {
                int ja = pair_j(s, &s->vcoul, j1);
                if (ja == ia) continue;
                double qj = q[ja];
                int qch_j = (qj != 0.0);
                const double *muj = mu[ja];
                int dip_j = (muj[0] != 0.0) || (muj[1] != 0.0) || (muj[2] != 0.0);
                int jta = s->typ[ja];
                double r[3];
                min_image(&s->cell, fr[3 * ia] - fr[3 * ja], fr[3 * ia + 1] - fr[3 * ja + 1],
                          fr[3 * ia + 2] - fr[3 * ja + 2], r);
                double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                if (d2 > s->lrcutsq) continue;
                double qij = qi * qj;
                double d = sqrt(d2), d3 = d2 * d, d5 = d3 * d2;
                double dm1 = 1.0 / d, dm3 = dm1 / d2, dm5 = dm3 / d2, dm7 = dm5 / d2;
                int ldamp = (p->lpoldamping[ita][ita][jta] || p->lpoldamping[jta][ita][jta]) && lcd;
                double fd1 = 1.0, fd2 = 1.0, fdd1 = 0.0, fdd2 = 0.0;
                if (ldamp) {
                    tt_damp(s->E_TT, p->pol_damp_b[ita][ita][jta], p->pol_damp_c[ita][ita][jta], d, &fd1, &fdd1,
                            p->pol_damp_k[ita][ita][jta]);
                    tt_damp(s->E_TT, p->pol_damp_b[jta][ita][jta], p->pol_damp_c[jta][ita][jta], d, &fd2, &fdd2,
                            p->pol_damp_k[jta][ita][jta]);
                }
                /* screening functions ewald.c:524-532 */
                double expon = exp(-a2 * d2) / piroot;
                double F0 = erfc(a * d);
                double F1 = F0 + 2.0 * a * d * expon;
                double F2 = F1 + 4.0 * a3 * d3 * expon / 3.0;
                double F3 = F2 + 8.0 * a5 * d5 * expon / 15.0;
                /* damped counterparts ewald.c:536-541 */
                double F1d1 = -fd1 + 1.0, F1d2 = -fd2 + 1.0;
                double F2d1 = F1d1 + (d / 3.0) * fdd1, F2d2 = F1d2 + (d / 3.0) * fdd2;
                double T0 = dm1 * F0;
                double T1[3], T1D1[3], T1D2[3];
                for (int i = 0; i < 3; i++) {
                    double t = -r[i] * dm3;
                    T1D1[i] = t * F1d1;
                    T1D2[i] = t * F1d2;
                    T1[i] = t * F1;
                }
                double F2_dm5 = 3.0 * F2 * dm5, F2d1_dm5 = 3.0 * F2d1 * dm5, F2d2_dm5 = 3.0 * F2d2 * dm5;
                double F1_dm3 = F1 * dm3, F1d1_dm3 = F1d1 * dm3, F1d2_dm3 = F1d2 * dm3;
                double T2[3][3], T2D1[3][3], T2D2[3][3];
                for (int j = 0; j < 3; j++)
                    for (int k = 0; k < 3; k++) {
                        T2[j][k] = r[j] * r[k] * F2_dm5 - (j == k ? F1_dm3 : 0.0);
                        T2D1[j][k] = r[j] * r[k] * F2d1_dm5 - (j == k ? F1d1_dm3 : 0.0);
                        T2D2[j][k] = r[j] * r[k] * F2d2_dm5 - (j == k ? F1d2_dm3 : 0.0);
                    }
                double T3[3][3][3];
                if (ld) {
                    double F3_dm7 = F3 * dm7 * 15.0;
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++)
                            for (int k = 0; k < 3; k++) {
                                double t = -r[i] * r[j] * r[k] * F3_dm7;
                                if (i == j) t += r[k] * F2_dm5;
                                if (i == k) t += r[j] * F2_dm5;
                                if (j == k) t += r[i] * F2_dm5;
                                T3[i][j][k] = t;
                            }
                }
                double fij[3] = {0.0, 0.0, 0.0};
                double *efi = A->ef + 3 * ia, *efj = A->ef + 3 * ja;
                double *efgi = A->efg + 9 * ia, *efgj = A->efg + 9 * ja;
                /* charge-charge ewald.c:703-731 */
                if (lq && (qch_i || qch_j)) {
                    A->u += qij * T0;
                    for (int i = 0; i < 3; i++) {
                        if (do_ef) {
                            efi[i] += -qj * T1[i];
                            efj[i] += qi * T1[i];
                            if (ldamp) {
                                efi[i] += qj * T1D1[i];
                                efj[i] += -qi * T1D2[i];
                            }
                        }
                        if (do_f) fij[i] = qij * T1[i];
                        if (do_efg)
                            for (int j = 0; j < 3; j++) {
                                efgi[3 * i + j] += -qj * T2[i][j];
                                efgj[3 * i + j] += -qi * T2[i][j];
                            }
                    }
                }
                /* dipole-dipole ewald.c:737-761 */
                if (ld && dip_i && dip_j) {
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++) {
                            A->u += -mui[i] * T2[i][j] * muj[j];
                            if (do_ef) {
                                efi[i] += muj[j] * T2[i][j];
                                efj[i] += mui[j] * T2[i][j];
                            }
                            if (do_efg || do_f)
                                for (int k = 0; k < 3; k++) {
                                    efgi[3 * i + j] += muj[k] * T3[i][j][k];
                                    efgj[3 * i + j] += -mui[k] * T3[i][j][k];
                                    fij[k] += -mui[i] * T3[i][k][j] * muj[j];
                                }
                        }
                }
                /* charge-dipole ewald.c:767-812 */
                if (lqd && (dip_i || dip_j)) {
                    for (int i = 0; i < 3; i++) {
                        A->u += -qi * T1[i] * muj[i];
                        A->u += qj * T1[i] * mui[i];
                        if (do_f)
                            for (int j = 0; j < 3; j++) {
                                fij[j] += -qi * T2[j][i] * muj[i];
                                fij[j] += qj * T2[j][i] * mui[i];
                            }
                        if (ldamp) {
                            A->u += qi * T1D2[i] * muj[i];
                            A->u += -qj * T1D1[i] * mui[i];
                            if (do_f)
                                for (int j = 0; j < 3; j++) {
                                    fij[j] += -qj * T2D1[i][j] * mui[i];
                                    fij[j] += qi * T2D2[i][j] * muj[i];
                                }
                        }
                    }
                }
                if (do_f) {
                    for (int k = 0; k < 3; k++) {
                        A->f[3 * ia + k] += -fij[k];
                        A->f[3 * ja + k] += fij[k];
                    }
                }
                if (do_s) {
                    A->tt[0] += -(r[0] * fij[0] + r[0] * fij[0]);
                    A->tt[1] += -(r[0] * fij[1] + r[1] * fij[0]);
                    A->tt[2] += -(r[0] * fij[2] + r[2] * fij[0]);
                    A->tt[3] += -(r[1] * fij[1] + r[1] * fij[1]);
                    A->tt[4] += -(r[1] * fij[2] + r[2] * fij[1]);
                    A->tt[5] += -(r[2] * fij[2] + r[2] * fij[2]);
                }
            }
        }
    }
    double uu = 0.0, tt[6] = {0};
    for (int t = 0; t < nt; t++) {
        uu += acc[t].u;
        for (int k = 0; k < 6; k++) tt[k] += acc[t].tt[k];
        for (int i = 0; i < 3 * n; i++) { f_dir[i] += acc[t].f[i]; ef_dir[i] += acc[t].ef[i]; }
        for (int i = 0; i < 9 * n; i++) efg_dir[i] += acc[t].efg[i];
        free(acc[t].f); free(acc[t].ef); free(acc[t].efg);
    }
    free(acc);
    free(fr);
    *u_dir = uu;
    if (do_s) {
        double full[3][3] = {{tt[0], tt[1], tt[2]}, {tt[1], tt[3], tt[4]}, {tt[2], tt[4], tt[5]}};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) tau_dir[i][j] = 0.5 * full[i][j] * s->cell.inveOmegaPU;
    }
}

/* multipole_ES_rec ewald.c:876-1051 with the charge densities of
 * kspace.c:189-302; do_efg is the caller's do_ef (ewald.c:169, Q6) and
 * only the upper triangle of the field gradient is accumulated
 * (ewald.c:966-973) */
static void ewald_rec(orc_sys *s, const double *q, const double (*mu)[3], int flags,
                      double *u_rec, double *ef_rec, double *efg_rec, double *f_rec, double tau_rec[3][3])
{
    const int lq = flags & ORC_LQ, ld = flags & ORC_LD;
    const int do_f = flags & ORC_FORCES, do_s = flags & ORC_STRESS;
    const int do_ef = flags & ORC_EF, do_efg = do_ef;
    int n = s->nion, nt = nthreads();
    struct_fact(s);
    double *F = (double *)calloc((size_t)nt * 3 * n, sizeof(double));
    double *EF = (double *)calloc((size_t)nt * 3 * n, sizeof(double));
    double *EFG = (double *)calloc((size_t)nt * 9 * n, sizeof(double));
    double *acc = (double *)calloc((size_t)nt * 8, sizeof(double));
#pragma omp parallel
    {
        int t = thread_id();
        double *Ft = F + (size_t)t * 3 * n, *Et = EF + (size_t)t * 3 * n, *Gt = EFG + (size_t)t * 9 * n;
        double *at = acc + 8 * t;
#pragma omp for schedule(dynamic, 8)
        for (int ik = 0; ik < s->nk; ik++) {
            if (s->kk[ik] == 0.0) continue;
            double kx = s->kx[ik], ky = s->ky[ik], kz = s->kz[ik], Ak = s->Ak[ik], kcoe = s->kcoe[ik];
            const double *c = s->ckr + (size_t)ik * n, *sn = s->skr + (size_t)ik * n;
            double rR = 0.0, rI = 0.0;
            for (int ia = 0; ia < n; ia++) {
                if (lq) {
                    rR += q[ia] * c[ia];
                    rI += q[ia] * sn[ia];
                }
                if (ld) {
                    double kmu = mu[ia][0] * kx + mu[ia][1] * ky + mu[ia][2] * kz;
                    rR += -kmu * sn[ia];
                    rI += kmu * c[ia];
                }
            }
            double str = (rR * rR + rI * rI) * Ak;
            at[0] += str;
            for (int ia = 0; ia < n; ia++) {
                double recarg = Ak * (rI * c[ia] - rR * sn[ia]);
                double recarg2 = Ak * (rR * c[ia] + rI * sn[ia]);
                double fx = kx * recarg, fy = ky * recarg, fz = kz * recarg;
                if (do_ef) {
                    Et[3 * ia] += -fx; Et[3 * ia + 1] += -fy; Et[3 * ia + 2] += -fz;
                }
                if (do_efg) {
                    Gt[9 * ia + 0] += kx * kx * recarg2;
                    Gt[9 * ia + 4] += ky * ky * recarg2;
                    Gt[9 * ia + 8] += kz * kz * recarg2;
                    Gt[9 * ia + 1] += kx * ky * recarg2;
                    Gt[9 * ia + 2] += kx * kz * recarg2;
                    Gt[9 * ia + 5] += ky * kz * recarg2;
                }
                if (do_f) {
                    double qi = q[ia];
                    Ft[3 * ia] += -qi * fx; Ft[3 * ia + 1] += -qi * fy; Ft[3 * ia + 2] += -qi * fz;
                    if (ld) {
                        double kmu = (mu[ia][0] * kx + mu[ia][1] * ky + mu[ia][2] * kz) * recarg2;
                        Ft[3 * ia] += kx * kmu; Ft[3 * ia + 1] += ky * kmu; Ft[3 * ia + 2] += kz * kmu;
                    }
                }
            }
            if (do_s) {
                at[1] += (1.0 - kcoe * kx * kx) * str;
                at[2] += -kcoe * kx * ky * str;
                at[3] += -kcoe * kx * kz * str;
                at[4] += (1.0 - kcoe * ky * ky) * str;
                at[5] += -kcoe * ky * kz * str;
                at[6] += (1.0 - kcoe * kz * kz) * str;
            }
        }
    }
    double uu = 0.0, tt[6] = {0};
    for (int t = 0; t < nt; t++) {
        uu += acc[8 * t];
        for (int k = 0; k < 6; k++) tt[k] += acc[8 * t + 1 + k];
        for (int i = 0; i < 3 * n; i++) { f_rec[i] += F[(size_t)t * 3 * n + i]; ef_rec[i] += EF[(size_t)t * 3 * n + i]; }
        for (int i = 0; i < 9 * n; i++) efg_rec[i] += EFG[(size_t)t * 9 * n + i];
    }
    free(F); free(EF); free(EFG); free(acc);
    double ttpiV = 2.0 * TPI * s->cell.inveOmega, tfpiV = 2.0 * ttpiV;
    *u_rec = uu * ttpiV;
    for (int i = 0; i < 3 * n; i++) { f_rec[i] *= tfpiV; ef_rec[i] *= tfpiV; }
    for (int i = 0; i < 9 * n; i++) efg_rec[i] *= tfpiV;
    double full[3][3] = {{tt[0], tt[1], tt[2]}, {tt[1], tt[3], tt[4]}, {tt[2], tt[4], tt[5]}};
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) tau_rec[j][k] = full[j][k] * ttpiV * s->cell.inveOmegaPU;
}

/* multipole_ES ewald.c:75-348 */
void orc_multipole_es(orc_sys *s, const double *q, const double *mu_, int flags,
                      double *u, double *pvir, double tau[9], double *ef, double *efg,
                      double *fx, double *fy, double *fz)
{
    const double (*mu)[3] = (const double(*)[3])mu_;
    int n = s->nion;
    double u_dir = 0.0, u_rec = 0.0;
    double tau_dir[3][3] = {{0}}, tau_rec[3][3] = {{0}};
    double *ef_dir = (double *)calloc((size_t)3 * n, sizeof(double));
    double *efg_dir = (double *)calloc((size_t)9 * n, sizeof(double));
    double *f_dir = (double *)calloc((size_t)3 * n, sizeof(double));
    double *ef_rec = (double *)calloc((size_t)3 * n, sizeof(double));
    double *efg_rec = (double *)calloc((size_t)9 * n, sizeof(double));
    double *f_rec = (double *)calloc((size_t)3 * n, sizeof(double));
    if (flags & ORC_DIR) ewald_dir(s, q, mu, flags, &u_dir, ef_dir, efg_dir, f_dir, tau_dir);
    if (flags & ORC_REC) ewald_rec(s, q, mu, flags, &u_rec, ef_rec, efg_rec, f_rec, tau_rec);
    /* self terms ewald.c:193-226 */
    double qsq = 0.0, musq = 0.0;
    for (int ia = 0; ia < n; ia++) {
        qsq += q[ia] * q[ia];
        musq += mu[ia][0] * mu[ia][0] + mu[ia][1] * mu[ia][1] + mu[ia][2] * mu[ia][2];
    }
    double a = s->alpha, a2 = a * a;
    double selfa = a / sqrt(PI);
    double selfa2 = 2.0 * selfa * a2 / 3.0;
    double u_self = -selfa * qsq - selfa2 * musq;
    *u = (u_dir + u_rec + u_self) * s->u.coul;
    for (int ia = 0; ia < n; ia++) {
        for (int i = 0; i < 3; i++) {
            ef[3 * ia + i] = 2.0 * selfa2 * mu[ia][i] + ef_dir[3 * ia + i] + ef_rec[3 * ia + i];
            for (int j = 0; j < 3; j++) {
                double self = (i == j) ? -2.0 * selfa2 * q[ia] : 0.0;
                efg[9 * ia + 3 * i + j] = self + efg_dir[9 * ia + 3 * i + j] + efg_rec[9 * ia + 3 * i + j];
            }
        }
        fx[ia] += f_dir[3 * ia] * s->u.coul + f_rec[3 * ia] * s->u.coul;
        fy[ia] += f_dir[3 * ia + 1] * s->u.coul + f_rec[3 * ia + 1] * s->u.coul;
        fz[ia] += f_dir[3 * ia + 2] * s->u.coul + f_rec[3 * ia + 2] * s->u.coul;
    }
    if (flags & ORC_STRESS)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) tau[3 * i + j] = (tau_dir[i][j] + tau_rec[i][j]) * s->u.coul;
    *pvir = 0.0;
    for (int i = 0; i < 3; i++) *pvir += tau[4 * i] * (1.0 / 3.0);
    free(ef_dir); free(efg_dir); free(f_dir); free(ef_rec); free(efg_rec); free(f_rec);
}

/* ------------------------------------------------------------------ PIM */
/* induced_moment pim.c:617-641: mu_i,a = sum_b pol[b][a] E_b on
 * polarisable ions only */
static void induced_moment(const orc_sys *s, double (*mu)[3], const double *ef)
{
    for (int ia = 0; ia < s->nion; ia++) {
        if (!s->p.lpolar[s->typ[ia]]) continue;
        for (int i = 0; i < 3; i++) {
            mu[ia][i] = 0.0;
            for (int j = 0; j < 3; j++) mu[ia][i] += s->pol[ia][j][i] * ef[3 * ia + j];
        }
    }
}

/* get_upol pim.c:145-161 */
static double get_upol(const orc_sys *s, const double (*mu)[3])
{
    double up = 0.0;
    for (int ia = 0; ia < s->nion; ia++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) up += mu[ia][i] * s->invpol[ia][i][j] * mu[ia][j];
    return up * 0.5;
}

/* get_rmsd_mu pim.c:431-449 */
static double get_rmsd(const orc_sys *s, const double (*mu)[3], const double *ef)
{
    double r = 0.0;
    int npol = 0;
    for (int ia = 0; ia < s->nion; ia++) {
        if (!s->p.lpolar[s->typ[ia]]) continue;
        npol++;
        for (int i = 0; i < 3; i++)
            if (s->pol[ia][i][i] != 0.0) {
                double e = mu[ia][i] / s->pol[ia][i][i] - ef[3 * ia + i];
                r += e * e;
            }
    }
    return sqrt(r / (double)npol);
}

/* ASPC coefficients extrapolate_dipole_aspc pim.c:335-379 */
static int aspc_coeffs(int ext_ord, double *B, double *W)
{
    switch (ext_ord) {
    case 0: B[0] = 1.0; *W = 0.0; return 0;
    case 1: B[0] = 2.0; B[1] = -1.0; *W = 2.0 / 3.0; return 0;
    case 2: B[0] = 2.5; B[1] = -2.0; B[2] = 0.5; *W = 0.6; return 0;
    case 3: B[0] = 2.8; B[1] = -2.8; B[2] = 1.2; B[3] = -0.2; *W = 4.0 / 7.0; return 0;
    case 4:
        B[0] = 3.0; B[1] = -24.0 / 7.0; B[2] = 27.0 / 14.0; B[3] = -4.0 / 7.0; B[4] = 1.0 / 14.0;
        *W = 5.0 / 9.0;
        return 0;
    case 5:
        B[0] = 22.0 / 7.0; B[1] = -55.0 / 14.0; B[2] = 55.0 / 21.0; B[3] = -22.0 / 21.0;
        B[4] = 5.0 / 21.0; B[5] = -1.0 / 42.0; *W = 6.0 / 11.0;
        return 0;
    default: return -1;
    }
}

/* order ramp pim.c:324-334 */
static int aspc_order(int istep, int order)
{
    if (istep >= order + 1) return order;
    return istep ? istep - 1 : 0;
}

/* predictor pim.c:382-400.  The reference raises ext_ord inside the ion
 * loop, so ions after the first also read coefficient slots that were not
 * set for this order; those slots multiply history entries that are still
 * zero during the ramp, so with finite slot contents every ion gets the
 * coefficients of the order computed at entry, which is what is used here. */
static void aspc_predict_ion(const double *B, int eo, int shift, double (*hist)[3], double *mu)
{
    for (int i = 0; i < 3; i++) {
        mu[i] = 0.0;
        for (int k = 0; k <= eo; k++) mu[i] += B[k] * hist[k][i];
    }
    for (int k = shift; k > 0; k--)
        for (int i = 0; i < 3; i++) hist[k][i] = hist[k - 1][i];
}

static void aspc_predict(orc_sys *s, double (*mu)[3], int istep)
{
    double B[MDFF_MAX_EXTRAPOLATE_ORDER + 1], W;
    int eo = aspc_order(istep, s->p.extrapolate_order);
    aspc_coeffs(eo, B, &W);
    int shift = (eo != s->p.extrapolate_order) ? eo + 1 : eo;
    for (int ia = 0; ia < s->nion; ia++) aspc_predict_ion(B, eo, shift, s->hist[ia], mu[ia]);
}

/* test hook: the predictor of one ion at step istep for extrapolate_order
 * `order`: hist[6][3] in (shifted on return), mu[3] out, and the corrector
 * weight W; returns the order used, -1 for an order without coefficients */
int orc_aspc_predict1(int istep, int order, double *hist, double *mu, double *W)
{
    double B[MDFF_MAX_EXTRAPOLATE_ORDER + 1];
    if (order < 0 || order > MDFF_MAX_EXTRAPOLATE_ORDER) return -1;
    int eo = aspc_order(istep, order);
    if (aspc_coeffs(eo, B, W)) return -1;
    aspc_predict_ion(B, eo, (eo != order) ? eo + 1 : eo, (double(*)[3])hist, mu);
    return eo;
}

/* corrector pim.c:403-420 */
static void aspc_correct(orc_sys *s, double (*mu)[3], const double *ef, int istep)
{
    double B[MDFF_MAX_EXTRAPOLATE_ORDER + 1], W;
    aspc_coeffs(aspc_order(istep, s->p.extrapolate_order), B, &W);
    double (*save)[3] = (double(*)[3])malloc(sizeof(double) * 3 * s->nion);
    memcpy(save, mu, sizeof(double) * 3 * s->nion);
    induced_moment(s, mu, ef);
    for (int ia = 0; ia < s->nion; ia++)
        for (int i = 0; i < 3; i++) mu[ia][i] = mu[ia][i] * W + (1.0 - W) * save[ia][i];
    free(save);
}

/* momentpolaSCF pim.c:177-306 */
static int moment_pola_scf(orc_sys *s, double (*mu_ind)[3], double *upol, int istep)
{
    int n = s->nion;
    double *ef = (double *)calloc((size_t)3 * n, sizeof(double));
    double *ef_stat = (double *)calloc((size_t)3 * n, sizeof(double));
    double *ef_ind = (double *)calloc((size_t)3 * n, sizeof(double));
    double *efg = (double *)calloc((size_t)9 * n, sizeof(double));
    double *zero = (double *)calloc((size_t)3 * n, sizeof(double));
    double *zq = (double *)calloc((size_t)n, sizeof(double));
    double u_stat, u_ind, pd, td[9] = {0};
    int fl = ORC_EF | ORC_DIR | ORC_REC;
    if (s->lqch) fl |= ORC_LQ;
    if (s->ldip) fl |= ORC_LD;
    if (s->ldmp) fl |= ORC_LDAMP;
    orc_multipole_es(s, s->q, (const double *)s->dip, fl, &u_stat, &pd, td, ef_stat, efg, zero, zero, zero);
    memcpy(ef, ef_stat, sizeof(double) * 3 * n);
    int iscf = 0;
    double rmsd = DBL_MAX, uupol = 0.0;
    while ((iscf < s->p.max_scf_pim_iter && rmsd > s->p.conv_tol_ind) || iscf < s->p.min_scf_pim_iter) {
        memset(mu_ind, 0, sizeof(double) * 3 * n);
        if (!iscf) aspc_predict(s, mu_ind, istep);
        else induced_moment(s, mu_ind, ef);
        uupol = get_upol(s, (const double(*)[3])mu_ind);
        orc_multipole_es(s, zq, (const double *)mu_ind, ORC_LD | ORC_EF | ORC_DIR | ORC_REC, &u_ind, &pd, td,
                         ef_ind, efg, zero, zero, zero);
        for (int i = 0; i < 3 * n; i++) ef[i] = ef_stat[i] + ef_ind[i];
        if (!iscf) aspc_correct(s, mu_ind, ef, istep);
        rmsd = get_rmsd(s, (const double(*)[3])mu_ind, ef);
        iscf++;
    }
    for (int ia = 0; ia < n; ia++)
        for (int i = 0; i < 3; i++) s->hist[ia][0][i] = mu_ind[ia][i];
    *upol = uupol * s->u.coul;
    s->th.n_scf = iscf;
    s->th.rmsd = rmsd;
    free(ef); free(ef_stat); free(ef_ind); free(efg); free(zero); free(zq);
    return (rmsd > s->p.conv_tol_ind) ? MDFF_ERR_SCF : MDFF_OK;
}

/* induced_moment_inner_kO pim.c:647-748: dmu from the external field
 * f_ext by omega-relaxed Jacobi sweeps on the field of dmu alone, real space
 * only with alphaES set to 0.001 for the sweeps (the self term of
 * multipole_ES ewald.c:193-226 at that alpha included); at most 21 sweeps,
 * until the rmsd of a sweep is below a tenth of the rmsd of f_ext */
static void induced_moment_inner_ko(orc_sys *s, double (*dmu)[3], const double *f_ext)
{
    int n = s->nion;
    double *ef_tot = (double *)malloc(sizeof(double) * 3 * n);
    double *ef_ind = (double *)calloc((size_t)3 * n, sizeof(double));
    double *efg = (double *)calloc((size_t)9 * n, sizeof(double));
    double *zero = (double *)calloc((size_t)3 * n, sizeof(double));
    double *zq = (double *)calloc((size_t)n, sizeof(double));
    double (*prev)[3] = (double(*)[3])malloc(sizeof(double) * 3 * n);
    double u, pd, td[9] = {0};
    memcpy(ef_tot, f_ext, sizeof(double) * 3 * n);
    double rmsd_ext = get_rmsd(s, (const double(*)[3])zero, ef_tot);
    induced_moment(s, dmu, ef_tot);
    const double w = s->p.omegakO;
    for (int ia = 0; ia < n; ia++)
        for (int i = 0; i < 3; i++) dmu[ia][i] *= w;
    const double alpha_save = s->alpha;
    s->alpha = 0.001;   /* pim.c:693-694 (the computed value is overwritten) */
    int it = 0;
    double rmsd_inner = DBL_MAX;
    while (it <= 20 && rmsd_inner > 0.1 * rmsd_ext) {
        it++;
        memcpy(prev, dmu, sizeof(double) * 3 * n);
        orc_multipole_es(s, zq, (const double *)dmu, ORC_LD | ORC_EF | ORC_DIR, &u, &pd, td, ef_ind, efg, zero, zero,
                         zero);
        for (int i = 0; i < 3 * n; i++) ef_tot[i] = f_ext[i] + ef_ind[i];
        rmsd_inner = get_rmsd(s, (const double(*)[3])dmu, ef_tot);
        s->ko_inner++;
        induced_moment(s, dmu, ef_tot);
        for (int ia = 0; ia < n; ia++)
            for (int i = 0; i < 3; i++) dmu[ia][i] = (1.0 - w) * prev[ia][i] + w * dmu[ia][i];
    }
    s->alpha = alpha_save;
    free(ef_tot); free(ef_ind); free(efg); free(zero); free(zq); free(prev);
}

/* momentpolaSCFkO pim.c:457-603: the SCF loop of momentpolaSCF with the
 * update mu += dmu from induced_moment_inner_kO on the residual field
 * f = E - mu diag(pol^-1); the ASPC predictor only when istep > 1 */
static int moment_pola_scf_ko(orc_sys *s, double (*mu_ind)[3], double *upol, int istep)
{
    int n = s->nion;
    double *ef = (double *)calloc((size_t)3 * n, sizeof(double));
    double *ef_stat = (double *)calloc((size_t)3 * n, sizeof(double));
    double *ef_ind = (double *)calloc((size_t)3 * n, sizeof(double));
    double *f_ind = (double *)calloc((size_t)3 * n, sizeof(double));
    double *efg = (double *)calloc((size_t)9 * n, sizeof(double));
    double *zero = (double *)calloc((size_t)3 * n, sizeof(double));
    double *zq = (double *)calloc((size_t)n, sizeof(double));
    double (*dmu)[3] = (double(*)[3])calloc((size_t)3 * n, sizeof(double));
    double u_stat, u_ind, pd, td[9] = {0};
    int fl = ORC_EF | ORC_DIR | ORC_REC;
    if (s->lqch) fl |= ORC_LQ;
    if (s->ldip) fl |= ORC_LD;
    if (s->ldmp) fl |= ORC_LDAMP;
    orc_multipole_es(s, s->q, (const double *)s->dip, fl, &u_stat, &pd, td, ef_stat, efg, zero, zero, zero);
    memcpy(ef, ef_stat, sizeof(double) * 3 * n);
    memcpy(f_ind, ef_stat, sizeof(double) * 3 * n);
    s->ko_inner = 0;
    int iscf = 0;
    double rmsd = DBL_MAX, uupol = 0.0;
    while ((iscf < s->p.max_scf_pim_iter && rmsd > s->p.conv_tol_ind) || iscf < s->p.min_scf_pim_iter) {
        if (!iscf && istep > 1) {
            aspc_predict(s, mu_ind, istep);
        } else {
            induced_moment_inner_ko(s, dmu, f_ind);
            for (int ia = 0; ia < n; ia++)
                for (int i = 0; i < 3; i++) mu_ind[ia][i] += dmu[ia][i];
        }
        uupol = get_upol(s, (const double(*)[3])mu_ind);
        orc_multipole_es(s, zq, (const double *)mu_ind, ORC_LD | ORC_EF | ORC_DIR | ORC_REC, &u_ind, &pd, td,
                         ef_ind, efg, zero, zero, zero);
        for (int i = 0; i < 3 * n; i++) ef[i] = ef_stat[i] + ef_ind[i];
        if (!iscf) aspc_correct(s, mu_ind, ef, istep);
        rmsd = get_rmsd(s, (const double(*)[3])mu_ind, ef);
        for (int ia = 0; ia < n; ia++) {
            if (!s->p.lpolar[s->typ[ia]]) continue;
            for (int i = 0; i < 3; i++) f_ind[3 * ia + i] = ef[3 * ia + i] - mu_ind[ia][i] * s->invpol[ia][i][i];
        }
        iscf++;
    }
    for (int ia = 0; ia < n; ia++)
        for (int i = 0; i < 3; i++) s->hist[ia][0][i] = mu_ind[ia][i];
    *upol = uupol * s->u.coul;
    s->th.n_scf = iscf;
    s->th.rmsd = rmsd;
    free(ef); free(ef_stat); free(ef_ind); free(f_ind); free(efg); free(zero); free(zq); free(dmu);
    return (rmsd > s->p.conv_tol_ind) ? MDFF_ERR_SCF : MDFF_OK;
}

/* get_dipoles coulombic.c:67-130 (algo scf or scfkO) */
static int get_dipoles(orc_sys *s, double (*mu)[3], double *upol, int istep)
{
    int rc = MDFF_OK;
    for (int ia = 0; ia < s->nion; ia++)
        for (int i = 0; i < 3; i++) mu[ia][i] = s->dip[ia][i];
    if (s->lpol) {
        if (s->p.algo_pim != MDFF_PIM_SCF && s->p.algo_pim != MDFF_PIM_SCFKO) return MDFF_ERR_UNSUPPORTED;
        double (*mi)[3] = (double(*)[3])calloc((size_t)3 * s->nion, sizeof(double));
        rc = s->p.algo_pim == MDFF_PIM_SCFKO ? moment_pola_scf_ko(s, mi, upol, istep)
                                             : moment_pola_scf(s, mi, upol, istep);
        for (int ia = 0; ia < s->nion; ia++)
            for (int i = 0; i < 3; i++) mu[ia][i] += mi[ia][i];
        s->ldip = 1; /* coulombic.c:115 */
        free(mi);
    }
    return rc;
}

/* engforce field.c:352-407 */
int orc_engforce(orc_sys *s, int istep)
{
    int n = s->nion, rc = MDFF_OK;
    s->istep = istep;
    for (int i = 0; i < n; i++) s->fx[i] = s->fy[i] = s->fz[i] = 0.0;
    if (s->p.lnmlj) engforce_nmlj(s, &s->th.u_lj, &s->th.pvir_lj, s->th.tau_lj);
    if (s->p.lbhmft || s->p.lbhmftd) engforce_bhmftd(s, &s->th.u_bhmftd, &s->th.pvir_bhmftd, s->th.tau_bhmftd);
    if (s->p.lcoulombic) {
        double *ef = (double *)calloc((size_t)3 * n, sizeof(double));
        double *efg = (double *)calloc((size_t)9 * n, sizeof(double));
        s->th.n_scf = 0;
        rc = get_dipoles(s, s->mu, &s->th.u_pol, istep);
        int fl = ORC_FORCES | ORC_STRESS | ORC_DIR | ORC_REC;
        if (s->lqch) fl |= ORC_LQ;
        if (s->ldip) fl |= ORC_LD;
        if (s->ldmp) fl |= ORC_LDAMP;
        orc_multipole_es(s, s->q, (const double *)s->mu, fl, &s->th.u_coul, &s->th.pvir_coul,
                         (double *)s->th.tau_coul, ef, efg, s->fx, s->fy, s->fz);
        free(ef); free(efg);
    }
    return rc;
}

/* --------------------------------------------------------- integrators */
static double calc_kin(const orc_sys *s)
{
    double k = 0.0;
    for (int i = 0; i < s->nion; i++)
        k += (s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i] + s->vz[i] * s->vz[i]) * s->mass[i];
    return k * 0.5;
}

/* prop_velocity_verlet integration.c:22-63 */
static int prop_vv(orc_sys *s, int istep)
{
    int n = s->nion;
    double dtsq2 = s->dt * s->dt * 0.5, dt2 = s->dt * 0.5;
    double *fs = (double *)malloc(sizeof(double) * 3 * n);
    for (int i = 0; i < n; i++) {
        fs[3 * i] = s->fx[i]; fs[3 * i + 1] = s->fy[i]; fs[3 * i + 2] = s->fz[i];
        s->rx[i] += s->vx[i] * s->dt + (s->fx[i] * dtsq2) * s->invmass[i];
        s->ry[i] += s->vy[i] * s->dt + (s->fy[i] * dtsq2) * s->invmass[i];
        s->rz[i] += s->vz[i] * s->dt + (s->fz[i] * dtsq2) * s->invmass[i];
    }
    s->pos_version++;
    int rc = orc_engforce(s, istep);
    for (int i = 0; i < n; i++) {
        s->vx[i] += (fs[3 * i] + s->fx[i]) * dt2 * s->invmass[i];
        s->vy[i] += (fs[3 * i + 1] + s->fy[i]) * dt2 * s->invmass[i];
        s->vz[i] += (fs[3 * i + 2] + s->fz[i]) * dt2 * s->invmass[i];
    }
    s->th.e_kin = calc_kin(s);
    free(fs);
    return rc;
}

/* prop_pos_vel_verlet integration.c:66-90 */
static int prop_pos_vel(orc_sys *s, double *kin, int istep)
{
    int n = s->nion;
    double dt2 = s->dt * 0.5;
    for (int i = 0; i < n; i++) {
        s->vx[i] += s->fx[i] * dt2 * s->invmass[i];
        s->vy[i] += s->fy[i] * dt2 * s->invmass[i];
        s->vz[i] += s->fz[i] * dt2 * s->invmass[i];
        s->rx[i] += s->vx[i] * s->dt;
        s->ry[i] += s->vy[i] * s->dt;
        s->rz[i] += s->vz[i] * s->dt;
    }
    s->pos_version++;
    int rc = orc_engforce(s, istep);
    for (int i = 0; i < n; i++) {
        s->vx[i] += s->fx[i] * dt2 * s->invmass[i];
        s->vy[i] += s->fy[i] * dt2 * s->invmass[i];
        s->vz[i] += s->fz[i] * dt2 * s->invmass[i];
    }
    *kin = calc_kin(s);
    return rc;
}

/* prop_leap_frog integration.c:93-137 (r(t-dt) starts at zero as
 * allocated in config.c:68) */
static int prop_lf(orc_sys *s, int istep)
{
    int rc = orc_engforce(s, istep);
    int n = s->nion;
    double dtsq = s->dt * s->dt, idt = 0.5 / s->dt;
    for (int i = 0; i < n; i++) {
        double ux = 2.0 * s->rx[i] - s->rxs[i] + s->fx[i] * dtsq * s->invmass[i];
        double uy = 2.0 * s->ry[i] - s->rys[i] + s->fy[i] * dtsq * s->invmass[i];
        double uz = 2.0 * s->rz[i] - s->rzs[i] + s->fz[i] * dtsq * s->invmass[i];
        s->vx[i] = idt * (ux - s->rxs[i]);
        s->vy[i] = idt * (uy - s->rys[i]);
        s->vz[i] = idt * (uz - s->rzs[i]);
        s->rxs[i] = s->rx[i]; s->rys[i] = s->ry[i]; s->rzs[i] = s->rz[i];
        s->rx[i] = ux; s->ry[i] = uy; s->rz[i] = uz;
    }
    s->pos_version++;
    s->th.e_kin = calc_kin(s);
    return rc;
}

/* chain_nhcn integration.c:189-294 (the forward sweep leaves vxi[n-2]
 * unscaled, integration.c:269-277) */
static void chain_nhcn(orc_sys *s, double *kin, const double *Q, double L)
{
    int nh = s->p.nhc_n, yo = s->p.nhc_yosh_order;
    double w[7] = {0};
    switch (yo) {
    case 1: w[0] = 1.0; break;
    case 3:
        w[0] = 1.0 / (2.0 - pow(2.0, 1.0 / 3.0));
        w[1] = 1.0 - 2.0 * w[0];
        w[2] = w[0];
        break;
    case 5:
        w[0] = 1.0 / (4.0 - pow(4, 1.0 / 3.0));
        w[1] = w[2] = w[3] = w[0];
        w[4] = 1.0 - 4.0 * w[0];
        break;
    case 7:
        w[0] = .78451361047756; w[1] = .235573213359357; w[2] = -1.17767998417887;
        w[3] = 1.0 - 2.0 * (w[0] + w[1] + w[2]);
        w[4] = w[2]; w[5] = w[1]; w[6] = w[0];
        break;
    default: break; /* 9: zero weights (integration.c:234-235) */
    }
    double *G = (double *)calloc(nh, sizeof(double));
    double s_ = 1.0;
    double T = s->kT;
    double *vxi = s->vxi, *xi = s->xi;
    G[0] = 2.0 * (*kin) - L * T;
    for (int k = 0; k < s->p.nhc_mults; k++) {
        for (int j = 0; j < yo; j++) {
            double dts = (yo <= 7 ? w[j] : 0.0) * s->dt / ((double)s->p.nhc_mults);
            double dts2 = dts * 0.5, dts4 = dts2 * 0.5, dts8 = dts4 * 0.5;
            G[nh - 1] = (vxi[nh - 2] * vxi[nh - 2] / Q[nh - 2] - T);
            vxi[nh - 1] += G[nh - 1] * dts4;
            for (int ih = nh - 2; ih > -1; ih--) {
                vxi[ih] *= exp(-vxi[ih + 1] * dts8 / Q[ih + 1]);
                vxi[ih] += G[ih] * dts4;
                vxi[ih] *= exp(-vxi[ih + 1] * dts8 / Q[ih + 1]);
            }
            s_ *= exp(-vxi[0] * dts2 / Q[0]);
            for (int i = 0; i < nh; i++) xi[i] += vxi[i] * dts2 / Q[i];
            G[0] = 2.0 * (*kin) * s_ * s_ - L * T;
            for (int ih = 0; ih < nh - 2; ih++) {
                vxi[ih] *= exp(-vxi[ih + 1] * dts8 / Q[ih + 1]);
                vxi[ih] += G[ih] * dts4;
                vxi[ih] *= exp(-vxi[ih + 1] * dts8 / Q[ih + 1]);
                G[ih + 1] = (vxi[ih] * vxi[ih] / Q[ih] - T);
            }
            vxi[nh - 1] += G[nh - 1] * dts4;
        }
    }
    *kin = 0.0;
    for (int i = 0; i < s->nion; i++) {
        s->vx[i] *= s_; s->vy[i] *= s_; s->vz[i] *= s_;
        *kin += (s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i] + s->vz[i] * s->vz[i]) * s->mass[i];
    }
    *kin *= 0.5;
    free(G);
}

/* nhcn integration.c:140-182 */
static int nhcn(orc_sys *s, int istep)
{
    int nh = s->p.nhc_n;
    double L = 3.0 * (double)s->nion;
    double *Q = (double *)malloc(sizeof(double) * nh);
    for (int i = 0; i < nh; i++) Q[i] = s->timesca * s->timesca * s->kT;
    Q[0] *= L;
    double kin = calc_kin(s);
    chain_nhcn(s, &kin, Q, L);
    int rc = prop_pos_vel(s, &kin, istep);
    chain_nhcn(s, &kin, Q, L);
    s->th.e_kin = calc_kin(s);
    double e = L * s->kT * s->xi[0] + s->vxi[0] * s->vxi[0] * 0.5 / Q[0];
    for (int i = 1; i < nh; i++) e += s->vxi[i] * s->vxi[i] * 0.5 / Q[i] + s->kT * s->xi[i];
    s->th.e_nvt = e;
    free(Q);
    return rc;
}

/* rescale_velocities kinetic.c:83-136 */
static void rescale_velocities(orc_sys *s)
{
    double ekin = calc_kin(s);
    double T = (2.0 * ekin) / (3.0 * (double)s->nion * s->u.boltz);
    double lambda;
    if (s->integ != MDFF_NVT_NHCN)
        lambda = sqrt(1.0 + (s->dt / s->tauT) * ((s->kT / T / s->u.boltz) - 1.0));
    else
        lambda = sqrt(1.0 + ((s->kT / T / s->u.boltz) - 1.0));
    double sx = 0, sy = 0, sz = 0;
    for (int i = 0; i < s->nion; i++) {
        s->vx[i] *= lambda; s->vy[i] *= lambda; s->vz[i] *= lambda;
        sx += s->vx[i]; sy += s->vy[i]; sz += s->vz[i];
    }
    double on = 1.0 / (double)s->nion;
    sx *= on; sy *= on; sz *= on;
    for (int i = 0; i < s->nion; i++) { s->vx[i] -= sx; s->vy[i] -= sy; s->vz[i] -= sz; }
}

/* init_velocities kinetic.c:15-28 (box_muller rand.c:20-32) */
void orc_init_velocities(orc_sys *s)
{
    srand(1);
    double sx = 0, sy = 0, sz = 0;
    for (int i = 0; i < s->nion; i++) {
        double rt = sqrt(s->kT * s->invmass[i]);
        double *v[3] = {&s->vx[i], &s->vy[i], &s->vz[i]};
        for (int c = 0; c < 3; c++) {
            double U = (double)rand() / (double)RAND_MAX, V = (double)rand() / (double)RAND_MAX;
            while (U == 0.0 || V == 0.0) {
                U = (double)rand() / (double)RAND_MAX;
                V = (double)rand() / (double)RAND_MAX;
            }
            *v[c] = rt * (0.0 + sqrt(-2.0 * log(U)) * cos(2.0 * PI * V) * 1.0);
        }
        sx += s->vx[i]; sy += s->vy[i]; sz += s->vz[i];
    }
    double on = 1.0 / (double)s->nion;
    sx *= on; sy *= on; sz *= on;
    for (int i = 0; i < s->nion; i++) { s->vx[i] += -sx; s->vy[i] += -sy; s->vz[i] += -sz; }
    rescale_velocities(s);
}

/* thermo.c:12-33 derived observables */
static void finish_thermo(orc_sys *s)
{
    mdff_thermo *t = &s->th;
    double l = 3.0 * (double)s->nion;
    t->temp_r = (2.0 * t->e_kin) / (l * s->u.boltz);
    double u_vdw = t->u_lj + t->u_bhmftd;
    t->u_tot = u_vdw + t->u_coul + t->u_pol;
    t->e_tot = t->u_tot + t->e_kin;
    t->h_tot = t->e_tot + t->e_nvt;
    double pv = t->pvir_lj + t->pvir_bhmftd + t->pvir_coul;
    t->pressure = pv + t->temp_r * s->u.boltz * s->cell.inveOmegaPU;
}

/* run_md md.c:196-211: kinetic energy and the force call at t=0 */
int orc_md_init(orc_sys *s)
{
    s->th.e_kin = calc_kin(s);
    int rc = orc_engforce(s, 0);
    s->th.istep = 0;
    finish_thermo(s);
    return rc;
}

/* one pass of the run_md loop md.c:229-248 */
int orc_md_step(orc_sys *s, int istep)
{
    int rc;
    switch (s->integ) {
    case MDFF_NVE_LF: rc = prop_lf(s, istep); break;
    case MDFF_NVT_NHCN: rc = nhcn(s, istep); break;
    default: rc = prop_vv(s, istep); break;
    }
    if (s->p.lverletL) orc_check_verletlist(s);
    if (s->rescale_allowed && istep < s->p.nequil && s->p.nequilT > 0 && istep % s->p.nequilT == 0)
        rescale_velocities(s);
    s->th.istep = istep;
    finish_thermo(s);
    return rc;
}

/* --------------------------------------------------------------- set-up */
/* init_field field.c:314-335, info_field field.c:186-220, init_config
 * config.c:124-198, check_md md.c:120-141 */
orc_sys *orc_create(const mdff_params *p, int nion, const int *typia, const double *A9,
                    const double *rx, const double *ry, const double *rz)
{
    if (nion < 1 || p->ntype < 1 || p->ntype > NT) return NULL;
    /* chain_nhcn exits on other Yoshida orders (integration.c:206-208) */
    if (p->integrator == MDFF_NVT_NHCN &&
        (p->nhc_n < 2 || p->nhc_mults < 1 ||
         !(p->nhc_yosh_order == 1 || p->nhc_yosh_order == 3 || p->nhc_yosh_order == 5 ||
           p->nhc_yosh_order == 7 || p->nhc_yosh_order == 9)))
        return NULL;
    orc_sys *s = (orc_sys *)calloc(1, sizeof(orc_sys));
    s->p = *p;
    s->nion = nion;
    s->ntype = p->ntype;
    units_init(&s->u, p->lreduced);
    cell_init(&s->cell, A9, s->u.press);
    size_t N = (size_t)nion;
    s->typ = (int *)malloc(sizeof(int) * N);
    memcpy(s->typ, typia, sizeof(int) * N);
    double **arr[12] = {&s->rx, &s->ry, &s->rz, &s->vx, &s->vy, &s->vz, &s->fx, &s->fy, &s->fz, &s->rxs, &s->rys, &s->rzs};
    for (int k = 0; k < 12; k++) *arr[k] = (double *)calloc(N, sizeof(double));
    memcpy(s->rx, rx, sizeof(double) * N);
    memcpy(s->ry, ry, sizeof(double) * N);
    memcpy(s->rz, rz, sizeof(double) * N);
    s->mass = (double *)malloc(sizeof(double) * N);
    s->invmass = (double *)malloc(sizeof(double) * N);
    s->q = (double *)calloc(N, sizeof(double));
    s->dip = (double(*)[3])calloc(N * 3, sizeof(double));
    s->pol = (double(*)[3][3])calloc(N * 9, sizeof(double));
    s->invpol = (double(*)[3][3])calloc(N * 9, sizeof(double));
    s->hist = (double(*)[MDFF_MAX_EXTRAPOLATE_ORDER + 1][3])calloc(N * 3 * (MDFF_MAX_EXTRAPOLATE_ORDER + 1), sizeof(double));
    s->mu = (double(*)[3])calloc(N * 3, sizeof(double));
    tt_init(s->E_TT);
    /* flags, field.c:195-219 */
    for (int it = 0; it < s->ntype; it++) {
        if (p->qit[it] != 0.0) s->lqch = 1;
        for (int j = 0; j < 3; j++) {
            if (p->dipit[it][j] != 0.0) s->ldip = 1;
            for (int k = 0; k < 3; k++) {
                if (p->quadit[it][j][k] != 0.0) s->lqua = 1;
                if (p->polit[it][j][k] != 0.0) s->lpol = 1;
            }
        }
        if (p->lpolar[it]) s->lpim = 1;
    }
    for (int t = 0; t < s->ntype; t++)
        for (int i = 0; i < s->ntype; i++)
            for (int k = 0; k < s->ntype; k++)
                if (p->lpoldamping[t][i][k]) s->ldmp = 1;
    s->lnonbonded = p->lnmlj || p->lbhmft || p->lbhmftd;
    s->srcutsq = p->cutshortrange * p->cutshortrange;
    s->lrcutsq = p->cutlongrange * p->cutlongrange;
    s->alpha = p->alphaES;
    memcpy(s->kES, p->kES, sizeof(s->kES));
    if (p->lcoulombic && p->lautoES) {
        /* set_autoES ewald.c:48-71 */
        double rc = p->cutlongrange, wmin = s->cell.w[0];
        for (int i = 1; i < 3; i++) if (s->cell.w[i] < wmin) wmin = s->cell.w[i];
        if (rc < 0.5 * wmin) rc = 0.5 * wmin;
        double eps = fabs(p->epsw) < 0.5 ? fabs(p->epsw) : 0.5;
        double tol = sqrt(fabs(log(eps * rc)));
        double a = sqrt((double)abs((int)(log(eps * rc * tol)))) / rc;
        double tol1 = sqrt(-log(eps * rc * (2.0 * tol * a) * (2.0 * tol * a)));
        s->alpha = a;
        for (int k = 0; k < 3; k++) s->kES[k] = (int)nearbyint(0.25 + s->cell.Anorm[k] * a * tol1 / PI);
    }
    /* per ion, config.c:141-190 */
    for (int ia = 0; ia < nion; ia++) {
        int it = s->typ[ia];
        s->mass[ia] = p->massit[it];
        s->invmass[ia] = 1.0 / p->massit[it];
        if (p->lcoulombic) {
            s->q[ia] = p->qit[it];
            for (int i = 0; i < 3; i++) s->dip[ia][i] = p->dipit[it][i];
        }
        if (p->lcoulombic && p->lpolar[it]) {
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) s->pol[ia][i][j] = p->polit[it][i][j];
            /* invertmatrix3x3 inverse_matrix.c:6-23 (LAPACK getrf/getri);
             * closed-form adjugate inverse here */
            double (*m)[3] = s->pol[ia];
            double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                         m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
            double id = 1.0 / det;
            s->invpol[ia][0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
            s->invpol[ia][0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
            s->invpol[ia][0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
            s->invpol[ia][1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
            s->invpol[ia][1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
            s->invpol[ia][1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
            s->invpol[ia][2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
            s->invpol[ia][2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
            s->invpol[ia][2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
        }
    }
    if (p->lcoulombic) kmesh_init(s);
    /* Verlet lists, config.c:72-78 */
    s->vnb.cut = p->cutshortrange + p->skindiff;
    s->vcoul.cut = p->cutlongrange + p->skindiff;
    s->vnb.point = (int *)calloc(N + 1, sizeof(int));
    s->vcoul.point = (int *)calloc(N + 1, sizeof(int));
    s->xvl = (double *)calloc(N, sizeof(double));
    s->yvl = (double *)calloc(N, sizeof(double));
    s->zvl = (double *)calloc(N, sizeof(double));
    /* md units, check_md md.c:120-141 */
    s->integ = (p->integrator == MDFF_NVE_LF && p->nequil > 0) ? MDFF_NVE_VV : p->integrator;
    s->rescale_allowed = (s->integ == MDFF_NVE_VV);
    double tauT = p->tauTberendsen;
    if (p->nequil > 0 && tauT == 0.0) tauT = p->dt;
    s->kT = p->temp * s->u.boltz;
    s->dt = p->dt * s->u.time;
    s->tauT = tauT * s->u.time;
    s->timesca = p->timesca_thermo * s->u.time;
    if (s->integ == MDFF_NVT_NHCN) {
        int nh = p->nhc_n > 0 ? p->nhc_n : 1;
        s->vxi = (double *)calloc(nh, sizeof(double));
        s->xi = (double *)calloc(nh, sizeof(double));
    }
    s->pos_version = 1;
    if (p->lverletL) orc_check_verletlist(s);
    return s;
}

void orc_destroy(orc_sys *s)
{
    if (!s) return;
    free(s->typ);
    free(s->rx); free(s->ry); free(s->rz); free(s->vx); free(s->vy); free(s->vz);
    free(s->fx); free(s->fy); free(s->fz); free(s->rxs); free(s->rys); free(s->rzs);
    free(s->mass); free(s->invmass); free(s->q); free(s->dip); free(s->pol); free(s->invpol);
    free(s->hist); free(s->mu);
    free(s->kx); free(s->ky); free(s->kz); free(s->kk); free(s->Ak); free(s->kcoe);
    free(s->ckr); free(s->skr);
    free(s->vnb.point); free(s->vnb.list); free(s->vcoul.point); free(s->vcoul.list);
    free(s->xvl); free(s->yvl); free(s->zvl);
    free(s->vxi); free(s->xi);
    free(s);
}

int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

void orc_set_positions(orc_sys *s, const double *rx, const double *ry, const double *rz)
{
    memcpy(s->rx, rx, sizeof(double) * s->nion);
    memcpy(s->ry, ry, sizeof(double) * s->nion);
    memcpy(s->rz, rz, sizeof(double) * s->nion);
    s->pos_version++;
}

void orc_set_velocities(orc_sys *s, const double *vx, const double *vy, const double *vz)
{
    memcpy(s->vx, vx, sizeof(double) * s->nion);
    memcpy(s->vy, vy, sizeof(double) * s->nion);
    memcpy(s->vz, vz, sizeof(double) * s->nion);
}

void orc_get_state(const orc_sys *s, double *rx, double *ry, double *rz,
                   double *vx, double *vy, double *vz, double *fx, double *fy, double *fz)
{
    size_t b = sizeof(double) * s->nion;
    if (rx) memcpy(rx, s->rx, b);
    if (ry) memcpy(ry, s->ry, b);
    if (rz) memcpy(rz, s->rz, b);
    if (vx) memcpy(vx, s->vx, b);
    if (vy) memcpy(vy, s->vy, b);
    if (vz) memcpy(vz, s->vz, b);
    if (fx) memcpy(fx, s->fx, b);
    if (fy) memcpy(fy, s->fy, b);
    if (fz) memcpy(fz, s->fz, b);
}

void orc_get_dipoles(const orc_sys *s, double *mu) { memcpy(mu, s->mu, sizeof(double) * 3 * s->nion); }

void orc_get_thermo(const orc_sys *s, mdff_thermo *th)
{
    *th = s->th;
    orc_sys *m = (orc_sys *)s;
    finish_thermo(m);
    *th = m->th;
    th->rmsd_fp64 = -1.0;   /* the oracle is FP64 throughout */
}

int orc_ko_inner(const orc_sys *s) { return s->ko_inner; }

void orc_get_ewald(const orc_sys *s, double *alpha, int *kES, int *nk)
{
    if (alpha) *alpha = s->alpha;
    if (kES) memcpy(kES, s->kES, sizeof(int) * 3);
    if (nk) *nk = s->nk;
}

/* ------------------------------------------------------- radial distribution */
/* default_rdf radial_distribution.c:37-42 (resg 0.1; the widths of lattice
 * cell.c:43-47) */
void orc_rdf_defaults(const double *A9, double *cutgr, int *nbins)
{
    cell_t c;
    cell_init(&c, A9, 1.0);
    double w = c.w[0];
    if (c.w[1] < w) w = c.w[1];
    if (c.w[2] < w) w = c.w[2];
    *cutgr = 0.5 * w;
    *nbins = (int)(*cutgr / 0.1) + 1;
}

/* get_gr radial_distribution.c:180-231: kardir of every ion
 * (cell.c:125-143), then for every ordered pair the fractional difference,
 * its minimum image in cartesian (pbc.c:7-14), and within cutgr the bin
 * (int)(r / resg); dirkar (cell.c:146-161) at the end */
long long orc_rdf_frame(int nion, int ntype, const int *typia, const double *A9, double *rx, double *ry,
                        double *rz, double cutgr, double resg, int nbins, long long *counts)
{
    cell_t c;
    cell_init(&c, A9, 1.0);
    const double cut2 = cutgr * cutgr;
    for (int i = 0; i < nion; i++) {
        double s[3];
        frac(&c, rx[i], ry[i], rz[i], s);
        rx[i] = s[0];
        ry[i] = s[1];
        rz[i] = s[2];
    }
    long long out = 0;
    for (int ia = 0; ia < nion; ia++) {
        const int ita = typia[ia];
        for (int ja = 0; ja < nion; ja++) {
            if (ja == ia) continue;
            double d[3];
            min_image(&c, rx[ia] - rx[ja], ry[ia] - ry[ja], rz[ia] - rz[ja], d);
            const double rijsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (rijsq < cut2) {
                const int igr = (int)(sqrt(rijsq) / resg);
                if (igr < 0 || igr >= nbins) {
                    out++;
                    continue;
                }
                counts[((size_t)igr * ntype + ita) * ntype + typia[ja]]++;
            }
        }
    }
    for (int i = 0; i < nion; i++) {
        const double x = rx[i], y = ry[i], z = rz[i];
        rx[i] = x * c.A[0][0] + y * c.A[1][0] + z * c.A[2][0];
        ry[i] = x * c.A[0][1] + y * c.A[1][1] + z * c.A[2][1];
        rz[i] = x * c.A[0][2] + y * c.A[1][2] + z * c.A[2][2];
    }
    return out;
}

/* write_rdf radial_distribution.c:120-141: the shell volume
 * 4 pi (resg r^2 + resg^3 / 12) / Omega at the bin centre and the count
 * divided by nconf vol n_it1 n_it2, in the reference's order of operations */
void orc_rdf_normalise(int nbins, int ntype, const long long *gr, int nconf, double resg, double omega,
                       const int *nionit, double *r, double *g)
{
    for (int igr = 0; igr < nbins; igr++) {
        const double rr = ((double)igr + 0.5) * resg;
        double vol = 4 * PI * (resg * rr * rr + (pow(resg, 3)) / 12);
        vol = vol / omega;
        r[igr] = rr;
        for (int a = 0; a < ntype; a++)
            for (int b = 0; b < ntype; b++) {
                const size_t k = ((size_t)igr * ntype + a) * ntype + b;
                g[k] = ((double)gr[k]) / (double)(nconf * vol * nionit[a] * nionit[b]);
            }
    }
}
