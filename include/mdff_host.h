/*
 * mdff_host.h -- host-side C front end of the engine: control-file parser,
 * POSFF/CONTFF/TRAJFF I/O, lattice helpers and the initial-velocity
 * generator.  Pure C, no device calls; exported by libmdff_gpu.so.
 *
 * Each entry point replaces a reference routine (cited per function) with a
 * reentrant one that fills caller-owned structs instead of globals and
 * returns a status instead of calling exit().
 */
#ifndef MDFF_HOST_H
#define MDFF_HOST_H

#include <stddef.h>
#include "mdff_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Unit factors (constants.c:8-45).  lreduced sets all of them to 1. */
typedef struct mdff_units {
    double boltz_unit, time_unit, press_unit, rho_unit, coul_unit;
} mdff_units;

/* Lattice derived quantities (cell.h:6-20, computed as cell.c:13-91). */
typedef struct mdff_cell {
    double A[3][3];      /* direct basis, rows a,b,c                       */
    double B[3][3];      /* reciprocal basis / Omega (cell.c:33-48)        */
    double Omega, inveOmega, inveOmegaPU;
    double Anorm[3], Bnorm[3], w[3];
} mdff_cell;

/* Defaults of every module (global.c:83-94, md.c:100-118, pim.c:72-85,
 * nmlj.c:110-120).  Replaces default_global/default_md/default_pim. */
void mdff_params_default(mdff_params *p);

/* Parse a control file (replaces read_global global.c:13, read_md md.c:29,
 * read_field field.c:31, read_bhmftd bhmftd.c:26, read_nmlj nmlj.c:39,
 * read_nonbonded nonbonded.c:7, read_pim pim.c:28).  ntype must come from
 * the POSFF (pair tables are read as ntype*(ntype+1)/2 values).  Defaults
 * are applied first.  Table symmetrisation as in bhmftd.c:98-110,
 * nmlj.c:205-212, field.c:209-219.  Returns MDFF_ERR_CONFIG on values the
 * reference rejects with exit(-1) (check_string tools.c:56-68). */
int mdff_read_control(const char *path, int ntype, mdff_params *p);

/* Units for a parameter set (gen_constants constants.c:8, reduced_units
 * constants.c:38 when lreduced). */
void mdff_units_get(const mdff_params *p, mdff_units *u);

/* lattice(CELL*) cell.c:13-91 for basis A (rows). press_unit enters
 * inveOmegaPU. */
void mdff_lattice(const double A[3][3], double press_unit, mdff_cell *c);

/* POSFF reader (read_config config.c:347-446): cartesian or Direct input,
 * optional velocity/force columns per Fposff, then the wrap to the
 * primary cell kardir -> s-nint(s) -> dirkar (config.c:442-444).
 * Types are assigned by count order (config.c:135-148). */
int mdff_read_posff(const char *path, int Fposff, mdff_config *cfg);

/* CONTFF/TRAJFF writer (write_config config.c:291-345): wraps a copy of
 * the positions, writes "%19.12f" columns; mode "w" or "a". */
int mdff_write_config(const char *path, const char *mode, const mdff_config *cfg);

int mdff_config_alloc(mdff_config *cfg, int nion);
void mdff_config_free(mdff_config *cfg);

/* Maxwell-Boltzmann velocities from the C library rand() stream exactly as
 * init_velocities kinetic.c:15-28 (maxwellboltzmann_velocities :42-64,
 * box_muller rand.c:20-32, rescale_velocities(0) :83-136).  The reference
 * never seeds rand() (mdff20.c:57), so reseed=1 calls srand(1) first to
 * reproduce a fresh process.  temp_K in Kelvin; mass per ion in amu. */
int mdff_host_init_velocities(const mdff_params *p, int nion, const double *mass,
                         double *vx, double *vy, double *vz, int reseed);

/* Berendsen / NHC velocity rescale (rescale_velocities kinetic.c:83-136). */
void mdff_host_rescale_velocities(const mdff_params *p, int nion, const double *mass,
                             double *vx, double *vy, double *vz);

/* Ewald parameters from epsw (set_autoES ewald.c:48-71), including the
 * integer abs() truncation and the cutoff raise that only feeds alpha/K.
 * Writes alpha and kES; returns the (possibly raised) cutlongrange. */
double mdff_auto_ewald(const mdff_params *p, const mdff_cell *cell,
                       double *alpha, int kES[3]);

/* sizeof of the shared structs, for bindings that mirror them */
size_t mdff_sizeof_params(void);
size_t mdff_sizeof_thermo(void);
size_t mdff_sizeof_config(void);

#ifdef __cplusplus
}
#endif
#endif /* MDFF_HOST_H */
