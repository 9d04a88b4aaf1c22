/*
 * mdff_batch.h -- C ABI of the batched engine (libmdff_gpu.so): many
 * independent periodic configurations of uneven size, one force-field
 * parameter set, evaluated together on one GPU, plus the hand-written
 * adjoint of a force-matching loss.  Plain C types only.
 *
 * The reference has no batched or differentiable entry point; each
 * configuration of a batch is evaluated exactly as engforce() evaluates a
 * static configuration (field.c:352-407: engforce_bhmftd_pbc / nmlj,
 * get_dipoles -> momentpolaSCF from zero dipoles, multipole_ES), with its
 * own cell, Ewald alpha, k box and cutoffs.  Every phase of a force call is
 * one launch over the whole batch (per run of configurations with the same
 * k-box tile count); the reciprocal-space sums are matrix-core GEMMs over
 * each configuration's k box, one (configuration, tile) work item per wave,
 * largest first, with tile shapes chosen by mdff_batch_autotune (k boxes
 * wider than 16 in x take the vector-ALU sums).
 *
 * Loss of mdff_batch_adjoint: L = sum over configurations and ions of
 * |F_i - F_ref,i|^2 (forces in eV/A).  Its gradient with respect to the
 * BHMFTD A, B, C, D, the polarisability tensors and the Tang-Toennies
 * charge-dipole damping b, c is formed by one implicit adjoint solve with the
 * symmetric dipole tensor (no stored SCF iterations):
 *   (pol^-1 - T) lambda = D_g E,  g = 2 (F - F_ref),
 * D_g E the derivative of the field at the polarisable ions along the ion
 * displacement g, then dL/dtheta = dPhi/dtheta - lambda . dR/dtheta with
 * Phi = sum g.F at fixed dipoles and R = coul (pol^-1 mu - E).
 */
#ifndef MDFF_BATCH_H
#define MDFF_BATCH_H

#include <stddef.h>

#include "mdff_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdff_batch mdff_batch;

/* Ewald and cutoff settings of one configuration */
typedef struct {
    double alpha;       /* Ewald alpha, 1/A (alphaES)                          */
    int kES[3];         /* k box |n_b| <= kES[b] (kES)                          */
    double rcut_long;   /* real-space Coulomb cutoff, A (cutlongrange)          */
    double rcut_short;  /* short-range pair cutoff, A (cutshortrange)           */
} mdff_batch_ewald;

/* gradient of the force-matching loss.  BHMFTD entries are the derivative
 * with respect to the shared value of the symmetric pair entry (both [i][j]
 * and [j][i] hold it); dpol[t] is the derivative with respect to the
 * entries of type t's polarisability tensor (polit); ddamp_b/c[t][s] with
 * respect to the damping of the field at a type-t ion from a type-s charge,
 * pol_damp_b/c[t][t][s] of the control file after its symmetrisation. */
typedef struct {
    double loss;
    double dA[MDFF_NTYPEMAX][MDFF_NTYPEMAX], dB[MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    double dC[MDFF_NTYPEMAX][MDFF_NTYPEMAX], dD[MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    double dpol[MDFF_NTYPEMAX][3][3];
    double ddamp_b[MDFF_NTYPEMAX][MDFF_NTYPEMAX], ddamp_c[MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    int adjoint_iterations;   /* Jacobi sweeps of the adjoint solve (max over configurations) */
    double adjoint_residual;  /* final relative residual (max over configurations) */
} mdff_grad;

/* Create a batch of nconf configurations: nion[c] ions each, types and
 * cartesian positions concatenated in configuration order, lattices
 * A9[c][9] (rows a, b, c).  ew[c] gives each configuration's Ewald and
 * cutoff settings; with ew == NULL every configuration takes alphaES, kES,
 * cutlongrange and cutshortrange of p (lautoES applied per cell).
 * precision: MDFF_FP64 or MDFF_MIXED (mdff_gpu.h).  Cutoffs must not exceed
 * half the smallest cell width (minimum image).  Static dipoles are not
 * supported (MDFF_ERR_UNSUPPORTED); static quadrupoles are accepted and, as
 * in the reference (whose Ewald sums read no quadrupole, ewald.c:646-688),
 * change no result. */
int mdff_batch_create(mdff_batch **out, const mdff_params *p, int nconf, const int *nion, const int *typia,
                      const double *A9, const double *rx, const double *ry, const double *rz,
                      const mdff_batch_ewald *ew, int precision, int device);
void mdff_batch_destroy(mdff_batch *b);

int mdff_batch_set_positions(mdff_batch *b, const double *rx, const double *ry, const double *rz);
/* new force-field parameters (same types, charges and polarisable set:
 * MDFF_ERR_CONFIG when they differ, or when no ion would stay polarisable
 * or one would become so); damping, pair potentials and the SCF solver
 * (conjugate gradients only for symmetric positive definite
 * polarisabilities) follow the new values.  The SCF of the next call
 * starts from the last converged dipoles. */
int mdff_batch_set_params(mdff_batch *b, const mdff_params *p);
/* zero the dipoles the next SCF starts from (the state of a fresh batch) */
int mdff_batch_reset_dipoles(mdff_batch *b);

/* One force call per configuration.  th[c] (may be NULL) receives the
 * energies, stress tensors, n_scf and rmsd of configuration c.  Returns
 * MDFF_ERR_SCF when some configuration stopped unconverged. */
int mdff_batch_engforce(mdff_batch *b, mdff_thermo *th);
int mdff_batch_get_forces(mdff_batch *b, double *fx, double *fy, double *fz);
int mdff_batch_get_dipoles(mdff_batch *b, double *mu);  /* [ntot][3] */

/* Force call followed by the adjoint: loss and gradient (FP64 batches
 * only).  fref_* are the target forces, concatenated like the ions. */
int mdff_batch_adjoint(mdff_batch *b, const double *fref_x, const double *fref_y, const double *fref_z,
                       mdff_grad *g);

/* Autotuner: times the candidate shapes of the batch kernels on this
 * batch and keeps the fastest.  out[4] = lanes per ion of the real-space
 * kernels, box points (matrix-core sums: 32) or k-vectors (vector sums) per
 * structure-factor item, ions per back-projection item (matrix-core: 16 or
 * 32; vector: 64, 128 or 256), persistent blocks per CU of the vector sums.
 * mdff_batch_set_tuning sets the same (0 keeps a value). */
int mdff_batch_autotune(mdff_batch *b, int out[4]);
int mdff_batch_set_tuning(mdff_batch *b, int lanes, int kpb, int ipb, int bpc);

/* sizeof(mdff_grad), for bindings that mirror the struct */
size_t mdff_batch_sizeof_grad(void);

/* Ewald settings in use for configuration c and its half-space k count */
int mdff_batch_get_ewald(mdff_batch *b, int c, double *alpha, int kES[3], int *nk);
/* device time of the last mdff_batch_engforce (ms) and total SCF sweeps */
int mdff_batch_get_timing(mdff_batch *b, double *ms, long long *sweeps);

#ifdef __cplusplus
}
#endif
#endif /* MDFF_BATCH_H */
