/*
 * posff.c -- lattice helpers, POSFF reader, CONTFF/TRAJFF writer, and the
 * Maxwell-Boltzmann initial velocities of the reference's MD start.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mdff_host.h"

#define MDFF_PI 3.1415926535897932385

/* cross product, expro tools.c:13-18 */
static void cross3(double h[3], const double *u1, const double *u2)
{
    h[0] = u1[1] * u2[2] - u1[2] * u2[1];
    h[1] = u1[2] * u2[0] - u1[0] * u2[2];
    h[2] = u1[0] * u2[1] - u1[1] * u2[0];
}

/* lattice() cell.c:13-91.  The volume is taken as sum_i B[i][0]*A[i][0]
 * (cell.c:38-41), exact for the orthorhombic cells the path is used with;
 * widths and norms are column norms as in cell.c:46-64. */
void mdff_lattice(const double A[3][3], double press_unit, mdff_cell *c)
{
    memset(c, 0, sizeof(*c));
    memcpy(c->A, A, sizeof(c->A));
    cross3(c->B[0], A[1], A[2]);
    cross3(c->B[1], A[2], A[0]);
    cross3(c->B[2], A[0], A[1]);
    double omega = c->B[0][0] * A[0][0] + c->B[1][0] * A[1][0] + c->B[2][0] * A[2][0];
    c->Omega = omega;
    c->inveOmega = 1.0 / omega;
    c->inveOmegaPU = 1.0 / (omega * press_unit);
    for (int i = 0; i < 3; i++)
        c->w[i] = omega / sqrt(c->B[0][i] * c->B[0][i] + c->B[1][i] * c->B[1][i] + c->B[2][i] * c->B[2][i]);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) c->B[i][j] = c->B[i][j] * c->inveOmega;
    for (int i = 0; i < 3; i++) {
        double a = 0.0, b = 0.0;
        for (int j = 0; j < 3; j++) {
            a += A[j][i] * A[j][i];
            b += c->B[j][i] * c->B[j][i];
        }
        c->Anorm[i] = sqrt(a);
        c->Bnorm[i] = sqrt(b);
    }
}

/* kardir cell.c:125-143 / dirkar cell.c:146-161 on one vector */
static void to_frac(const double B[3][3], double *x, double *y, double *z)
{
    double v1 = *x * B[0][0] + *y * B[0][1] + *z * B[0][2];
    double v2 = *x * B[1][0] + *y * B[1][1] + *z * B[1][2];
    double v3 = *x * B[2][0] + *y * B[2][1] + *z * B[2][2];
    *x = v1; *y = v2; *z = v3;
}

static void to_cart(const double A[3][3], double *x, double *y, double *z)
{
    double v1 = *x * A[0][0] + *y * A[1][0] + *z * A[2][0];
    double v2 = *x * A[0][1] + *y * A[1][1] + *z * A[2][1];
    double v3 = *x * A[0][2] + *y * A[1][2] + *z * A[2][2];
    *x = v1; *y = v2; *z = v3;
}

int mdff_config_alloc(mdff_config *cfg, int nion)
{
    cfg->nion = nion;
    size_t n = (size_t)(nion > 0 ? nion : 1);
    cfg->typia = (int *)calloc(n, sizeof(int));
    double **arr[9] = {&cfg->rx, &cfg->ry, &cfg->rz, &cfg->vx, &cfg->vy, &cfg->vz, &cfg->fx, &cfg->fy, &cfg->fz};
    for (int k = 0; k < 9; k++) *arr[k] = (double *)calloc(n, sizeof(double));
    for (int k = 0; k < 9; k++)
        if (!*arr[k]) return MDFF_ERR_NOMEM;
    return cfg->typia ? MDFF_OK : MDFF_ERR_NOMEM;
}

void mdff_config_free(mdff_config *cfg)
{
    free(cfg->typia);
    free(cfg->rx); free(cfg->ry); free(cfg->rz);
    free(cfg->vx); free(cfg->vy); free(cfg->vz);
    free(cfg->fx); free(cfg->fy); free(cfg->fz);
    cfg->typia = NULL;
    cfg->rx = cfg->ry = cfg->rz = cfg->vx = cfg->vy = cfg->vz = cfg->fx = cfg->fy = cfg->fz = NULL;
}

/* fscanf("%lf") on the next token; a token that is not a number is left
 * in place (the reader tolerates missing v/f columns, config.c:412-422) */
static int scan_double(FILE *fp, double *v)
{
    return fscanf(fp, "%lf", v) == 1;
}

int mdff_read_posff(const char *path, int Fposff, mdff_config *cfg)
{
    memset(cfg, 0, sizeof(*cfg));
    FILE *fp = fopen(path, "r");
    if (!fp) return MDFF_ERR_IO;
    int nion = 0;
    char buf[4096];
    if (fscanf(fp, "%d", &nion) != 1 || nion < 1) { fclose(fp); return MDFF_ERR_IO; }
    if (fscanf(fp, "%4095s", buf) != 1) { fclose(fp); return MDFF_ERR_IO; }
    strncpy(cfg->name, buf, MDFF_MAX_LEN);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (!scan_double(fp, &cfg->A[i][j])) { fclose(fp); return MDFF_ERR_IO; }
    if (fscanf(fp, "%d", &cfg->ntype) != 1 || cfg->ntype < 1 || cfg->ntype > MDFF_NTYPEMAX) {
        fclose(fp);
        return MDFF_ERR_IO;
    }
    for (int it = 0; it < cfg->ntype; it++) {
        if (fscanf(fp, "%4095s", buf) != 1) { fclose(fp); return MDFF_ERR_IO; }
        strncpy(cfg->atypit[it], buf, MDFF_MAX_LEN);
    }
    int sum = 0;
    for (int it = 0; it < cfg->ntype; it++) {
        if (fscanf(fp, "%d", &cfg->nionit[it]) != 1) { fclose(fp); return MDFF_ERR_IO; }
        sum += cfg->nionit[it];
    }
    int rc = mdff_config_alloc(cfg, nion);
    if (rc) { fclose(fp); return rc; }
    char cpos[4096];
    if (fscanf(fp, "%4095s", cpos) != 1) { fclose(fp); return MDFF_ERR_IO; }
    for (int ia = 0; ia < nion; ia++) {
        if (fscanf(fp, "%4095s ", buf) != 1) { fclose(fp); return MDFF_ERR_IO; }
        if (fscanf(fp, "%lf %lf %lf ", &cfg->rx[ia], &cfg->ry[ia], &cfg->rz[ia]) != 3) {
            fclose(fp);
            return MDFF_ERR_IO;
        }
        if (Fposff > 0) {
            int c = fscanf(fp, "%lf %lf %lf ", &cfg->vx[ia], &cfg->vy[ia], &cfg->vz[ia]);
            (void)c;
        }
        if (Fposff > 1) {
            int c = fscanf(fp, "%lf %lf %lf ", &cfg->fx[ia], &cfg->fy[ia], &cfg->fz[ia]);
            (void)c;
        }
    }
    fclose(fp);
    /* config_input_ok config.c:117-122 */
    if (sum != nion) return MDFF_ERR_CONFIG;
    /* types by count order, config.c:135-148 */
    int cc = 0;
    for (int it = 0; it < cfg->ntype; it++)
        for (int k = 0; k < cfg->nionit[it]; k++) cfg->typia[cc++] = it;
    mdff_cell cell;
    mdff_lattice((const double(*)[3])cfg->A, 1.0, &cell);
    int direct = (strcmp(cpos, "Direct") == 0) || (strcmp(cpos, "D") == 0);
    for (int ia = 0; ia < nion; ia++) {
        double x = cfg->rx[ia], y = cfg->ry[ia], z = cfg->rz[ia];
        if (direct) to_cart(cell.A, &x, &y, &z);
        /* config.c:442-444 */
        to_frac(cell.B, &x, &y, &z);
        x -= nearbyint(x);
        y -= nearbyint(y);
        z -= nearbyint(z);
        to_cart(cell.A, &x, &y, &z);
        cfg->rx[ia] = x; cfg->ry[ia] = y; cfg->rz[ia] = z;
    }
    return MDFF_OK;
}

int mdff_write_config(const char *path, const char *mode, const mdff_config *cfg)
{
    FILE *fp = fopen(path, mode);
    if (!fp) return MDFF_ERR_IO;
    mdff_cell cell;
    mdff_lattice((const double(*)[3])cfg->A, 1.0, &cell);
    fprintf(fp, "%d \n", cfg->nion);
    fprintf(fp, "%s \n", cfg->name);
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) fprintf(fp, "%19.12f ", cfg->A[i][j]);
        fprintf(fp, " \n");
    }
    fprintf(fp, "%d \n", cfg->ntype);
    for (int it = 0; it < cfg->ntype; it++) fprintf(fp, "%s ", cfg->atypit[it]);
    fprintf(fp, "\n");
    for (int it = 0; it < cfg->ntype; it++) fprintf(fp, "%d ", cfg->nionit[it]);
    fprintf(fp, " \n");
    fprintf(fp, "Cartesian \n");
    for (int ia = 0; ia < cfg->nion; ia++) {
        double x = cfg->rx[ia], y = cfg->ry[ia], z = cfg->rz[ia];
        to_frac(cell.B, &x, &y, &z);
        x -= nearbyint(x);
        y -= nearbyint(y);
        z -= nearbyint(z);
        to_cart(cell.A, &x, &y, &z);
        fprintf(fp, "%s ", cfg->atypit[cfg->typia[ia]]);
        fprintf(fp, "%19.12f %19.12f %19.12f ", x, y, z);
        fprintf(fp, "%19.12f %19.12f %19.12f ", cfg->vx[ia], cfg->vy[ia], cfg->vz[ia]);
        fprintf(fp, "%19.12f %19.12f %19.12f ", cfg->fx[ia], cfg->fy[ia], cfg->fz[ia]);
        fprintf(fp, "\n");
    }
    fclose(fp);
    return MDFF_OK;
}

/* randD rand.c:14-16 and box_muller rand.c:20-32 */
static double rand_unit(void) { return (double)rand() / (double)(RAND_MAX); }

static double box_muller(double mean, double sigma)
{
    double U = rand_unit(), V = rand_unit();
    while (U == 0.0 || V == 0.0) {
        U = rand_unit();
        V = rand_unit();
    }
    double R = sqrt(-2.0 * log(U));
    double C = cos(2.0 * MDFF_PI * V);
    return mean + R * C * sigma;
}

static double kinetic(int n, const double *m, const double *vx, const double *vy, const double *vz)
{
    double k = 0.0;
    for (int i = 0; i < n; i++) k += (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]) * m[i];
    return k * 0.5;
}

/* effective integrator after check_md md.c:128-131 (nve-lf with nequil>0
 * runs as nve-vv) */
static int effective_integrator(const mdff_params *p)
{
    return (p->integrator == MDFF_NVE_LF && p->nequil > 0) ? MDFF_NVE_VV : p->integrator;
}

void mdff_host_rescale_velocities(const mdff_params *p, int nion, const double *mass,
                             double *vx, double *vy, double *vz)
{
    mdff_units u;
    mdff_units_get(p, &u);
    double kT = p->temp * u.boltz_unit;
    double tauT = p->tauTberendsen;
    if (p->nequil > 0 && tauT == 0.0) tauT = p->dt; /* md.c:122 */
    double dt_i = p->dt * u.time_unit, tau_i = tauT * u.time_unit;
    double ekin = kinetic(nion, mass, vx, vy, vz);
    double T = (2.0 * ekin) / (3.0 * (double)nion * u.boltz_unit);
    double lambda;
    if (effective_integrator(p) != MDFF_NVT_NHCN)
        lambda = sqrt(1.0 + (dt_i / tau_i) * ((kT / T / u.boltz_unit) - 1.0));
    else
        lambda = sqrt(1.0 + ((kT / T / u.boltz_unit) - 1.0));
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int i = 0; i < nion; i++) {
        vx[i] *= lambda; vy[i] *= lambda; vz[i] *= lambda;
        sx += vx[i]; sy += vy[i]; sz += vz[i];
    }
    double on = 1.0 / (double)nion;
    sx *= on; sy *= on; sz *= on;
    for (int i = 0; i < nion; i++) {
        vx[i] -= sx; vy[i] -= sy; vz[i] -= sz;
    }
}

int mdff_host_init_velocities(const mdff_params *p, int nion, const double *mass,
                         double *vx, double *vy, double *vz, int reseed)
{
    if (nion < 1 || !mass) return MDFF_ERR_ARG;
    mdff_units u;
    mdff_units_get(p, &u);
    if (reseed) srand(1);
    double kT = p->temp * u.boltz_unit;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int i = 0; i < nion; i++) {
        double rt = sqrt(kT * (1.0 / mass[i]));
        vx[i] = rt * box_muller(0.0, 1.0);
        vy[i] = rt * box_muller(0.0, 1.0);
        vz[i] = rt * box_muller(0.0, 1.0);
        sx += vx[i]; sy += vy[i]; sz += vz[i];
    }
    double on = 1.0 / (double)nion;
    sx *= on; sy *= on; sz *= on;
    for (int i = 0; i < nion; i++) {
        vx[i] += -sx; vy[i] += -sy; vz[i] += -sz;
    }
    mdff_host_rescale_velocities(p, nion, mass, vx, vy, vz);
    return MDFF_OK;
}

/* set_autoES ewald.c:48-71 */
double mdff_auto_ewald(const mdff_params *p, const mdff_cell *cell, double *alpha, int kES[3])
{
    double rc = p->cutlongrange;
    double wmin = cell->w[0];
    for (int i = 1; i < 3; i++) if (cell->w[i] < wmin) wmin = cell->w[i];
    double rcut = 0.5 * wmin;
    if (rc < rcut) rc = rcut;
    double eps = fabs(p->epsw) < 0.5 ? fabs(p->epsw) : 0.5;
    double tol = sqrt(fabs(log(eps * rc)));
    /* abs() of an int: the double argument is truncated toward zero */
    double a = sqrt((double)abs((int)(log(eps * rc * tol)))) / rc;
    double tol1 = sqrt(-log(eps * rc * (2.0 * tol * a) * (2.0 * tol * a)));
    *alpha = a;
    for (int k = 0; k < 3; k++) kES[k] = (int)nearbyint(0.25 + cell->Anorm[k] * a * tol1 / MDFF_PI);
    return rc;
}
