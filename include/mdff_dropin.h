/*
 * mdff_dropin.h -- same-name drop-ins for mdff20's force path.
 *
 * mdff20_amd/dropin/mdff_dropin.c defines these functions with the
 * reference's own signatures.  It is compiled against the reference's
 * headers (-I<mdff20>/src, for its globals) and linked into the reference's
 * executable together with libmdff_gpu.so, in place of the definitions it
 * replaces (INTEGRATION.md gives the link line).  Every function reads and
 * writes the reference's globals exactly as the function it replaces does,
 * and runs the computation on the GPU through the handle API of mdff_gpu.h:
 * one device context, created on first use from the parsed globals
 * (control-file parameters, simuCell, typia, positions) and kept until exit.
 * Positions are uploaded on every call; the SCF history, neighbour list and
 * k-space tables stay on the device.
 *
 * Environment: MDFF_PRECISION=mixed selects the mixed form (default FP64),
 * MDFF_DEVICE=<n> the HIP device (default 0).  Errors print a message and
 * call exit(-1), as the reference does for configuration errors; an
 * unconverged SCF is accepted silently, as in the reference.
 */
#ifndef MDFF_DROPIN_H
#define MDFF_DROPIN_H

#include <stdbool.h>

#include "mdff_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* engforce field.h:36, field.c:352-407: fx/fy/fz, u_lj, u_bhmftd, u_coul,
 * u_pol, pvir_*, tau_*; dipia_ind[.][0] receives the induced dipoles and
 * ldip is set as coulombic.c:115 and field.c:397-401 set it */
void engforce(void);

/* multipole_ES ewald.h:21-24, ewald.c:75-348: u, pvir, tau, ef, efg; forces
 * (coulombic units times coul_unit) added to fx/fy/fz when do_forces */
void multipole_ES(double *q, double (*mu)[3], double (*theta)[3][3], double *u, double *pvir, double tau[3][3],
                  double (*ef)[3], double (*efg)[3][3], bool lq, bool ld, bool lt, bool ldamp, bool do_forces,
                  bool do_stress, bool do_ef, bool do_efg, bool do_dir, bool do_rec, bool inpim, bool update_sf);

/* get_dipoles coulombic.h:17, coulombic.c:67-130: static + induced dipoles */
void get_dipoles(double (*mu)[3], double *upol);

/* momentpolaSCF pim.h:26, pim.c:177-306: induced dipoles only */
void momentpolaSCF(double (*mu)[3], double *upol);

/* engforce_bhmftd_pbc bhmftd.h:22 and engforce_nmlj_pbc nmlj.h:29: energy,
 * virial and stress of the pair potential, forces added to fx/fy/fz */
void engforce_bhmftd_pbc(double *u, double *pvir, double tau[3][3]);
void engforce_nmlj_pbc(double *u, double *pvir, double tau[3][3]);

/* TT_damping_functions tt_damp.h:7 (host evaluation of the device formula) */
void TT_damping_functions(double b, double c, double r, double *f, double *fd, int order);

/* check_verletlist / gen_pbc_verletlist verlet.h:20-21: the device list is
 * checked (rebuilt when the skin criterion of verlet.c:124-170 asks) and
 * updatevl counts the rebuilds; the host lists are not filled */
void check_verletlist(void);
void gen_pbc_verletlist(void);

/* prop_agate integration.h:3, integration.c:298-333: one step of the
 * control file's integrator (nve-lf, nve-vv, nvt-nhcn) with its force call
 * on the device.  The positions and velocities of the globals are uploaded
 * first (the caller may have changed them, e.g. rescale_velocities), and
 * rx/ry/rz, vx/vy/vz, fx/fy/fz, e_kin, e_nvt and the force call's globals
 * come back.  The thermostat chain (vxi, xi) and the leap-frog r(t-dt)
 * stay on the device. */
void prop_agate(int step);

/* run_md md.h:48, md.c:193-282: the whole MD loop on the device.  Between
 * two steps with output the steps run as one device launch sequence (step,
 * Verlet check, rescale, md.c:235-248); at an output step the globals are
 * updated and the reference's own info_thermo / write_config write stdout,
 * OSZIFF, CONTFF and TRAJFF as md.c:250-277 does.  The induced-dipole
 * history stays on the device (dipia_ind[.][0] holds the last step's). */
void run_md(void);

/* get_gr radial_distribution.h:22, radial_distribution.c:180-231: the
 * histogram of the current frame (rx/ry/rz, typia, simuCell) on the device
 * (mdff_rdf.h), added into gr as the reference adds it (ordered pairs per
 * type pair, and every pair once more into gr[.][0][0]); rx/ry/rz are left
 * as its kardir/dirkar round trip leaves them.  nbins must hold every
 * distance below cutgr at resg (the reference writes past gr otherwise);
 * else an error message and exit(-1). */
void get_gr(void);

/* invertmatrix3x3 inverse_matrix.h, inverse_matrix.c:6-23 (LAPACK there):
 * the closed-form 3x3 inverse in place; init_config calls it for every
 * polarisable ion */
void invertmatrix3x3(double A[3][3]);

/* the engine parameters the drop-ins build from the reference's parsed
 * globals (what the device context is created from) */
int mdff_dropin_params(mdff_params *p);

/* release the device context (also registered with atexit on first use) */
void mdff_dropin_release(void);

#ifdef __cplusplus
}
#endif
#endif /* MDFF_DROPIN_H */
