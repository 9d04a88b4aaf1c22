/*
 * ref_driver.c -- fixture and timing driver for the reference build in
 * oracle/_ref (TEST INFRASTRUCTURE).  Linked against libmdff20_ref.so, which
 * holds the reference's own src/*.c compiled unchanged except for
 * inverse_matrix.c (needs lapacke.h, absent) and mdff20.c (its main).
 *
 * It repeats main()'s initialisation order (mdff20.c:59-92), the
 * reference's own init_config included, and then either one static force
 * call (mdff20.c:94-107), the run_md loop body (md.c:229-248) without its
 * file output, or run_md itself (md.c:193-282, which writes OSZIFF, CONTFF
 * and TRAJFF into the case directory), writing every observable at %.17g.
 * A polarisable input needs the LAPACK inverse: the plain reference build
 * refuses it (exit status 5); the drop-in build runs it.
 *
 * usage: mdff20_ref_driver <dir with control and POSFF> static
 *        mdff20_ref_driver <dir> md <nsteps> <state-dump-period>
 *        mdff20_ref_driver <dir> run_md
 *        mdff20_ref_driver <dir> time <n_induced_calls>
 *        mdff20_ref_driver <dir> rdf [write]
 *
 * The rdf mode follows main's RDF branch (mdff20.c:108-110): the steps of
 * init_rdf (radial_distribution.c:162-178) and rdf_readtraj (:75-99) over
 * the case's TRAJFF, writing the integer counts of gr and the positions get_gr
 * leaves; write_rdf (:101-150) runs only with `write` (its first bin indexes
 * grr with the still uninitialised mp, :128).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <omp.h>

#include "cell.h"
#include "config.h"
#include "constants.h"
#include "coulombic.h"
#include "ewald.h"
#include "field.h"
#include "global.h"
#include "integration.h"
#include "io.h"
#include "kinetic.h"
#include "kspace.h"
#include "md.h"
#include "mpi_mdff.h"
#include "thermo.h"
#include "verlet.h"
#include "bhmftd.h"
#include "nmlj.h"
#include "radial_distribution.h"

static FILE *out;

/* present when the driver is linked with the same-name drop-ins */
extern int mdff_dropin_params(void *p) __attribute__((weak));

/* The reference's own init_config (config.c:124-198).  For a polarisable
 * type it calls invertmatrix3x3 (inverse_matrix.c:6-23), which needs LAPACK:
 * absent here, so the reference build leaves that unit out and cannot run a
 * polarisable input.  Linked with the drop-ins, the drop-in
 * invertmatrix3x3 of libmdff_gpu.so takes its place.  The `time` mode only
 * times routines that never read the inverse (multipole_ES calls), so it
 * runs init_config with the polarisable flags cleared and restores them. */
static int ref_init_config(const char *mode)
{
    int anypol = 0;
    for (int it = 0; it < ntype; it++)
        if (lcoulombic && lpolar[it]) anypol = 1;
    if (!anypol || mdff_dropin_params) {
        init_config();
        return 0;
    }
    if (strcmp(mode, "time") != 0) {
        fprintf(stderr, "polarisable types: init_config needs invertmatrix3x3 (LAPACK), "
                        "absent from this build\n");
        return 5;
    }
    bool keep[NTYPEMAX];
    for (int it = 0; it < NTYPEMAX; it++) { keep[it] = lpolar[it]; lpolar[it] = false; }
    init_config();
    for (int it = 0; it < NTYPEMAX; it++) lpolar[it] = keep[it];
    return 0;
}

static void dump_vec(const char *key, const double *v, int n)
{
    fprintf(out, "%s %d", key, n);
    for (int i = 0; i < n; i++) fprintf(out, " %.17g", v[i]);
    fprintf(out, "\n");
}

static void dump_scalars(int step, int n_scf)
{
    fprintf(out, "step %d\n", step);
    fprintf(out, "n_scf %d\n", n_scf);
    fprintf(out, "verlet_updates %d\n", updatevl);
    double sc[10] = {e_kin, e_nvt, u_lj, u_bhmftd, u_coul, u_pol, pvir_lj, pvir_bhmftd, pvir_coul, 0.0};
    dump_vec("scalars", sc, 9);
    dump_vec("tau_lj", &tau_lj[0][0], 9);
    dump_vec("tau_bhmftd", &tau_bhmftd[0][0], 9);
    dump_vec("tau_coul", &tau_coul[0][0], 9);
}

static void dump_state(void)
{
    dump_vec("rx", rx, nion); dump_vec("ry", ry, nion); dump_vec("rz", rz, nion);
    dump_vec("vx", vx, nion); dump_vec("vy", vy, nion); dump_vec("vz", vz, nion);
    dump_vec("fx", fx, nion); dump_vec("fy", fy, nion); dump_vec("fz", fz, nion);
    if (lcoulombic) {
        double *mu = malloc(3 * nion * sizeof(double));
        for (int ia = 0; ia < nion; ia++)
            for (int i = 0; i < 3; i++)
                mu[3 * ia + i] = dipia[ia][i] + (lpim ? dipia_ind[ia][0][i] : 0.0);
        dump_vec("mu", mu, 3 * nion);
        free(mu);
    }
}

static int scf_calls(void) { return cc_multipole_ES_dir_pim; }

static double wall(void) { return omp_get_wtime(); }

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s dir static|md|time ...\n", argv[0]);
        return 2;
    }
    if (chdir(argv[1]) != 0) {
        perror("chdir");
        return 2;
    }
    const char *mode = argv[2];
    out = fopen("ref_out.txt", "w");
    if (!out) return 3;
    if (!freopen("ref_stdout.txt", "w", stdout)) return 3;
    myrank = 0;
    numprocs = 1;
    init_io();
    gen_constants();
    controlfn = "control";
    init_global(controlfn);
    read_config();
    init_md(controlfn);
    init_field(controlfn);
    int irc = ref_init_config(mode);
    if (irc) return irc;
    do_split(nion, numprocs, myrank, &atomDec, "atoms");
    do_split(kcoul.nk, numprocs, myrank, &kcoul.kptDec, "k-points");
    if (egrator == 2) {
        for (int i = 0; i < nhc_n; i++) { vxi[i] = 0.0; xi[i] = 0.0; }
    }
    fprintf(out, "alphaES %.17g\n", alphaES);
    fprintf(out, "kES %d %d %d\n", kES[0], kES[1], kES[2]);
    fprintf(out, "nk %d\n", kcoul.nk);
    if (!strcmp(mode, "params")) {
        /* the drop-ins' engine parameters, as raw bytes in hex */
        if (!mdff_dropin_params) {
            fprintf(stderr, "params: not linked with the drop-ins\n");
            return 2;
        }
        unsigned char buf[1 << 20];
        memset(buf, 0, sizeof(buf));
        if (mdff_dropin_params(buf) != 0) return 4;
        size_t nb = argc > 3 ? (size_t)atol(argv[3]) : 0;
        if (nb == 0 || nb > sizeof(buf)) return 2;
        fprintf(out, "params %zu ", nb);
        for (size_t i = 0; i < nb; i++) fprintf(out, "%02x", buf[i]);
        fprintf(out, "\n");
        fclose(out);
        return 0;
    }
    if (lverletL) check_verletlist();

    if (!strcmp(mode, "static")) {
        e_kin = calc_kin();
        int c0 = scf_calls();
        engforce();
        dump_scalars(0, lpim ? scf_calls() - c0 - 1 : 0);
        dump_state();
    } else if (!strcmp(mode, "md")) {
        int nsteps = argc > 3 ? atoi(argv[3]) : npas;
        int period = argc > 4 ? atoi(argv[4]) : nsteps;
        init_velocities();
        dump_vec("v0x", vx, nion); dump_vec("v0y", vy, nion); dump_vec("v0z", vz, nion);
        e_kin = calc_kin();
        int c0 = scf_calls();
        engforce();
        dump_scalars(0, lpim ? scf_calls() - c0 - 1 : 0);
        if (period > 0 && 0 % period == 0) dump_state();
        for (istep = 1; istep < nsteps + 1; istep++) {
            c0 = scf_calls();
            prop_agate(istep);
            int ns = lpim ? scf_calls() - c0 - 1 : 0;
            if (lverletL) check_verletlist();
            if (rescale_allowed && (istep < nequil) && (istep % nequilT == 0)) rescale_velocities(1);
            dump_scalars(istep, ns);
            if ((period > 0 && istep % period == 0) || istep == nsteps) dump_state();
        }
    } else if (!strcmp(mode, "run_md")) {
        /* main's MD branch (mdff20.c:88-92): velocities, then run_md */
        init_velocities();
        dump_vec("v0x", vx, nion); dump_vec("v0y", vy, nion); dump_vec("v0z", vz, nion);
        int c0 = scf_calls();
        run_md();
        /* n_scf: the dipole Ewald calls of the SCF over the whole run */
        dump_scalars(npas, lpim ? scf_calls() - c0 : 0);
        dump_state();
    } else if (!strcmp(mode, "time")) {
        /* one force call split into the pieces of a PIM step, each timed */
        int nind = argc > 3 ? atoi(argv[3]) : 4;
        double (*ef)[3] = malloc(nion * sizeof(*ef));
        double (*efg)[3][3] = malloc(nion * sizeof(*efg));
        double (*mu)[3] = malloc(nion * sizeof(*mu));
        double *zq = calloc(nion, sizeof(double));
        double u, pv, tau[3][3] = {{0}};
        for (int i = 0; i < nion; i++) fx[i] = fy[i] = fz[i] = 0.0;
        double t0 = wall();
        if (lbhmft || lbhmftd) engforce_bhmftd_pbc(&u_bhmftd, &pvir_bhmftd, tau_bhmftd);
        if (lnmlj) engforce_nmlj_pbc(&u_lj, &pvir_lj, tau_lj);
        double t1 = wall();
        multipole_ES(qia, dipia, quadia, &u, &pv, tau, ef, efg, lqch, ldip, lqua, ldmp,
                     false, false, true, false, true, true, true, true);
        double t2 = wall();
        for (int ia = 0; ia < nion; ia++)
            for (int i = 0; i < 3; i++) mu[ia][i] = lpolar[typia[ia]] ? polit[typia[ia]][i][i] * ef[ia][i] : 0.0;
        for (int k = 0; k < nind; k++)
            multipole_ES(zq, mu, quadia, &u, &pv, tau, ef, efg, false, true, false, false,
                         false, false, true, false, true, true, true, false);
        double t3 = wall();
        multipole_ES(qia, mu, quadia, &u_coul, &pvir_coul, tau_coul, ef, efg, lqch, true, lqua, ldmp,
                     true, true, false, false, true, true, false, !lpim);
        double t4 = wall();
        fprintf(out, "threads %d\n", omp_get_max_threads());
        fprintf(out, "t_short %.9g\n", t1 - t0);
        fprintf(out, "t_static %.9g\n", t2 - t1);
        fprintf(out, "t_induced %.9g\n", nind > 0 ? (t3 - t2) / nind : 0.0);
        fprintf(out, "t_force %.9g\n", t4 - t3);
        fprintf(out, "n_induced %d\n", nind);
    } else if (!strcmp(mode, "rdf")) {
        default_rdf();
        read_rdf(controlfn);
        alloc_rdf();
        FILE *fp = fopen("TRAJFF", "r");
        if (!fp) return 6;
        for (int ic = 0; ic < nconf; ic++) {
            read_next_traj(fp);
            get_gr();
        }
        fclose(fp);
        fprintf(out, "rdf_nbins %d\n", nbins);
        fprintf(out, "rdf_ntype %d\n", ntype);
        fprintf(out, "rdf_nconf %d\n", nconf);
        fprintf(out, "resg %.17g\n", resg);
        fprintf(out, "cutgr %.17g\n", cutgr);
        fprintf(out, "omega %.17g\n", simuCell.Omega);
        double *g = malloc(sizeof(double) * (size_t)nbins * ntype * ntype);
        for (int b = 0; b < nbins; b++)
            for (int i = 0; i < ntype; i++)
                for (int j = 0; j < ntype; j++) g[((size_t)b * ntype + i) * ntype + j] = (double)gr[b][i][j];
        dump_vec("gr", g, nbins * ntype * ntype);
        free(g);
        double ni[NTYPEMAX];
        for (int i = 0; i < ntype; i++) ni[i] = (double)nionit[i];
        dump_vec("nionit", ni, ntype);
        dump_vec("rx", rx, nion); dump_vec("ry", ry, nion); dump_vec("rz", rz, nion);
        if (argc > 3 && !strcmp(argv[3], "write")) write_rdf();
    } else {
        fprintf(stderr, "unknown mode %s\n", mode);
        return 2;
    }
    fclose(out);
    return 0;
}
