/*
 * mdff_oracle.h -- CPU FP64 restatement of mdff20's force and integration
 * path.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and the
 * bench's CPU-baseline leg load it as the checker; the product engine
 * (libmdff_gpu.so) never links or calls it.
 *
 * Every routine follows a reference routine, cited at its definition in
 * mdff_oracle.c.  State is held in an orc_sys handle instead of globals.
 */
#ifndef MDFF_ORACLE_H
#define MDFF_ORACLE_H

#include "../include/mdff_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys orc_sys;

/* flags of multipole_ES (ewald.h:21-24), packed as bits */
enum {
    ORC_LQ = 1 << 0, ORC_LD = 1 << 1, ORC_LT = 1 << 2, ORC_LDAMP = 1 << 3,
    ORC_FORCES = 1 << 4, ORC_STRESS = 1 << 5, ORC_EF = 1 << 6, ORC_EFG = 1 << 7,
    ORC_DIR = 1 << 8, ORC_REC = 1 << 9
};

orc_sys *orc_create(const mdff_params *p, int nion, const int *typia, const double *A9,
                    const double *rx, const double *ry, const double *rz);
void orc_destroy(orc_sys *s);
int orc_set_threads(int n);

void orc_set_positions(orc_sys *s, const double *rx, const double *ry, const double *rz);
void orc_set_velocities(orc_sys *s, const double *vx, const double *vy, const double *vz);
void orc_get_state(const orc_sys *s, double *rx, double *ry, double *rz,
                   double *vx, double *vy, double *vz, double *fx, double *fy, double *fz);
/* total dipoles (static + induced) of the last force call, [N][3] */
void orc_get_dipoles(const orc_sys *s, double *mu);
void orc_get_thermo(const orc_sys *s, mdff_thermo *th);
/* derived Ewald parameters (after lautoES) and k count */
void orc_get_ewald(const orc_sys *s, double *alpha, int *kES, int *nk);
/* scfkO: inner real-space sweeps of the last SCF (induced_moment_inner_kO) */
int orc_ko_inner(const orc_sys *s);

/* initial velocities of an MD start (kinetic.c:15-28) from rand() after
 * srand(1) */
void orc_init_velocities(orc_sys *s);

/* one force evaluation at the current positions, reference engforce
 * (field.c:352-407); istep is the step number the ASPC ramp sees */
int orc_engforce(orc_sys *s, int istep);
/* the Verlet list check run after every step (verlet.c:124-170) */
void orc_check_verletlist(orc_sys *s);

/* multipole_ES (ewald.c:75-348) on caller arrays; forces are ADDED to fx.. */
void orc_multipole_es(orc_sys *s, const double *q, const double *mu, int flags,
                      double *u, double *pvir, double tau[9], double *ef, double *efg,
                      double *fx, double *fy, double *fz);

/* test hook: the ASPC predictor of one ion (pim.c:314-400) at step istep
 * for extrapolate_order `order`; hist[6][3] in and shifted on return, mu[3]
 * out, *W the corrector weight; returns the order used or -1 */
int orc_aspc_predict1(int istep, int order, double *hist, double *mu, double *W);

/* run_md (md.c:193-282) without file output: md_init does the t=0 force
 * call; md_step advances one step (prop_agate + list check + rescale) */
int orc_md_init(orc_sys *s);
int orc_md_step(orc_sys *s, int istep);

/* radial distribution (radial_distribution.c).  orc_rdf_defaults: cutgr and
 * nbins of default_rdf (:37-42) for cell A9 (rows a, b, c).  orc_rdf_frame:
 * get_gr (:180-231) over one frame, ordered pairs ia != ja, counts ADDED to
 * counts[nbins][ntype][ntype] (without the reference's extra all-pair
 * increment of gr[.][0][0], :219); the positions are left as get_gr leaves
 * them (kardir then dirkar in place); returns the pairs whose bin fell
 * outside [0, nbins) (not counted).  orc_rdf_normalise: the normalisation of
 * write_rdf (:120-141) of counts gr[nbins][ntype][ntype] over nconf frames:
 * r[igr] and g[igr][it1][it2]. */
void orc_rdf_defaults(const double *A9, double *cutgr, int *nbins);
long long orc_rdf_frame(int nion, int ntype, const int *typia, const double *A9, double *rx, double *ry,
                        double *rz, double cutgr, double resg, int nbins, long long *counts);
void orc_rdf_normalise(int nbins, int ntype, const long long *gr, int nconf, double resg, double omega,
                       const int *nionit, double *r, double *g);

#ifdef __cplusplus
}
#endif
#endif
