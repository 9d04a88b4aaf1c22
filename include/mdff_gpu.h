/*
 * mdff_gpu.h -- C ABI of the MI355X engine for mdff20's force and
 * integration path (libmdff_gpu.so).  Plain C types only.
 *
 * Handle API: one context owns the device state of one periodic system
 * (positions, velocities, forces, dipoles, neighbour list, k-space tables,
 * thermostat) on one GPU and one HIP stream.  Every entry point returns an
 * mdff status code (mdff_types.h) and never exits the process.
 *
 * Reference interfaces replaced (FilipeVasconcelos/mdff20, src/):
 *   mdff_engforce        <- void engforce(void)                field.h:36, field.c:352-407
 *   mdff_multipole_es    <- void multipole_ES(...)             ewald.h:21-24, ewald.c:75-348
 *   mdff_get_dipoles     <- void get_dipoles(mu, upol)         coulombic.h:17, coulombic.c:67-130
 *   mdff_pair_short      <- engforce_bhmftd_pbc / engforce_nmlj_pbc  bhmftd.h:22, nmlj.h:29
 *   mdff_check_verlet    <- void check_verletlist(void)        verlet.h:21, verlet.c:124-170
 *   mdff_propagate       <- void prop_agate(int)               integration.h:3, integration.c:298-333
 *   mdff_md_step         <- prop_agate + the rest of the body of run_md's loop
 *                                                              md.c:229-248
 *   mdff_md_run          <- void run_md(void)                  md.h:48, md.c:193-282 (no file output)
 *   mdff_tt_damping      <- TT_damping_functions(b,c,r,f,fd,order)  tt_damp.h:7
 * The global-state same-name wrappers (engforce(), multipole_ES(), ...) that
 * a maintainer links into the reference's own main live in mdff_dropin.h.
 */
#ifndef MDFF_GPU_H
#define MDFF_GPU_H

#include <stddef.h>
#include "mdff_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdff_ctx mdff_ctx;

/* arithmetic forms */
enum {
    MDFF_FP64 = 0,  /* everything in FP64 (parity form)                       */
    MDFF_MIXED = 1  /* FP32 real space and FP32 S(k) tables/contraction with
                       FP64 accumulation of S(k), fields, forces; FP64 SCF
                       residual, energies, virial, self terms, NHC chain      */
};

/* multipole_ES flags (argument order of ewald.h:21-24), packed as bits */
enum {
    MDFF_ES_LQ = 1 << 0, MDFF_ES_LD = 1 << 1, MDFF_ES_LT = 1 << 2, MDFF_ES_LDAMP = 1 << 3,
    MDFF_ES_FORCES = 1 << 4, MDFF_ES_STRESS = 1 << 5, MDFF_ES_EF = 1 << 6, MDFF_ES_EFG = 1 << 7,
    MDFF_ES_DIR = 1 << 8, MDFF_ES_REC = 1 << 9, MDFF_ES_INPIM = 1 << 10, MDFF_ES_UPDATE_SF = 1 << 11
};

/* Library version string and the HIP device count (0 when no GPU). */
const char *mdff_version(void);
int mdff_device_count(void);

/* Create a context for nion ions of types typia[] (count-ordered as POSFF),
 * lattice A (row-major 3x3, rows a,b,c) and cartesian positions.  All
 * derived settings of init_field/init_config/check_md are applied
 * (field.c:186-335, config.c:124-198, md.c:120-141), including lautoES.
 * precision: MDFF_FP64 or MDFF_MIXED.  device: HIP device ordinal. */
int mdff_create(mdff_ctx **out, const mdff_params *p, int nion, const int *typia,
                const double *A9, const double *rx, const double *ry, const double *rz,
                int precision, int device);
void mdff_destroy(mdff_ctx *ctx);

/* state transfer (host <-> device, synchronous).  mdff_set_positions keeps
 * the Verlet list: after moving ions by more than half the skin, call
 * mdff_check_verlet before the next force call, as run_md checks the list
 * after every step (md.c:241). */
int mdff_set_positions(mdff_ctx *ctx, const double *rx, const double *ry, const double *rz);
int mdff_set_velocities(mdff_ctx *ctx, const double *vx, const double *vy, const double *vz);
int mdff_get_positions(mdff_ctx *ctx, double *rx, double *ry, double *rz);
int mdff_get_velocities(mdff_ctx *ctx, double *vx, double *vy, double *vz);
int mdff_get_forces(mdff_ctx *ctx, double *fx, double *fy, double *fz);
/* total dipoles (static + induced) of the last force call, [nion][3] */
int mdff_get_dipoles(mdff_ctx *ctx, double *mu);
/* electric field at the ions after the last SCF, [nion][3] (e/A^2) */
int mdff_get_field(mdff_ctx *ctx, double *ef);
int mdff_get_thermo(mdff_ctx *ctx, mdff_thermo *th);
/* Ewald parameters in use (after lautoES) and the half-space k count */
int mdff_get_ewald(mdff_ctx *ctx, double *alpha, int *kES, int *nk);

/* Initial velocities of an MD start: host rand() after srand(1), exactly as
 * init_velocities kinetic.c:15-28, uploaded to the device. */
int mdff_init_velocities(mdff_ctx *ctx);

/* Verlet list check/rebuild at the current positions (verlet.c:124-170).
 * The first call always builds. */
int mdff_check_verlet(mdff_ctx *ctx);

/* mdff_check_verlet without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (tight serving
 * loops, code that forbids hipMalloc).  It runs the same check and rebuild,
 * with its scratch (the check's block partials and count, the cell grid's
 * arrays) in one fixed array per GPU held until the process ends (about
 * 3 MiB), and never grows the row capacity: a list that needs more room
 * returns MDFF_ERR_LIST_OVERFLOW (call mdff_check_verlet once to grow it).
 * Limits: at most 131072 ions and 131072 grid cells, else MDFF_ERR_SCRATCH.
 * Calls on one GPU share that array: calls that overlap on a GPU race. */
int mdff_check_verlet_noalloc(mdff_ctx *ctx);

/* One force evaluation at the current positions (engforce field.c:352-407):
 * short-range pair forces, PIM dipoles (SCF), multipole Ewald forces,
 * energies and stress.  istep is the step number seen by the ASPC ramp.
 * Returns MDFF_ERR_SCF if the SCF stopped unconverged (results valid, as
 * the reference accepts them). */
int mdff_engforce(mdff_ctx *ctx, int istep, mdff_thermo *th);

/* Short-range pair part only (engforce_bhmftd_pbc / engforce_nmlj_pbc):
 * energy u, pvir and stress tau of the active potential; forces left in
 * the context force array. */
int mdff_pair_short(mdff_ctx *ctx, double *u, double *pvir, double tau[9]);

/* multipole_ES on caller-supplied host arrays of nion rows: q[n], mu[n][3]
 * (theta is accepted for signature parity and, as in the reference, takes
 * no part in any term), flags MDFF_ES_*.  Outputs: u, pvir, tau[9],
 * ef[n][3], efg[n][3][3]; forces (coulombic units) are ADDED to fx/fy/fz
 * when non-NULL, as the reference adds them to the global force arrays. */
int mdff_multipole_es(mdff_ctx *ctx, const double *q, const double *mu, const double *theta,
                      int flags, double *u, double *pvir, double tau[9], double *ef, double *efg,
                      double *fx, double *fy, double *fz);

/* get_dipoles: static + SCF-induced dipoles at the current positions;
 * mu[n][3] out, upol in eV. */
int mdff_get_dipoles_scf(mdff_ctx *ctx, int istep, double *mu, double *upol);

/* run_md start (md.c:196-211): kinetic energy + force call at t=0. */
int mdff_md_init(mdff_ctx *ctx, mdff_thermo *th);
/* prop_agate (integration.c:298-333): one step of the configured
 * integrator (nve-lf, nve-vv or nvt-nhcn) with its force call, nothing else
 * (no Verlet check, no rescale).  State stays on the device. */
int mdff_propagate(mdff_ctx *ctx, int istep, mdff_thermo *th);
/* One MD step (prop_agate + Verlet check + equilibration rescale,
 * md.c:235-248).  State stays on the device. */
int mdff_md_step(mdff_ctx *ctx, int istep, mdff_thermo *th);
/* nsteps steps from step first_step; thermo of the last step.  Host and
 * device synchronise only when the SCF loop needs its convergence flag. */
int mdff_md_run(mdff_ctx *ctx, int first_step, int nsteps, mdff_thermo *th);

/* Tang-Toennies damping f, f' at one distance (TT_damping_functions
 * tt_damp.c:16-32), host evaluation of the exact device formula. */
void mdff_tt_damping(double b, double c, double r, double *f, double *fd, int order);

/* Tuning: lanes per ion in the neighbour kernels and the CG steps' pair-tensor
 * field (8, 16, 32 or 64, default 16; 0 keeps),
 * reciprocal-space kernel family katoms (-1 vector ALU; 1 the matrix-core
 * sums, split-FP16 SCF steps in the mixed form, the default; 2 the same with
 * FP32 SCF steps; 4 as 1 with the forward's partial sums, the S(k) reduction
 * and the pair-tensor field of the CG steps in separate kernels instead of
 * fused; 0 keeps) and ksplit, the target
 * block count of the matrix-core kernels (0 keeps).  No counterpart in the
 * reference. */
int mdff_set_tuning(mdff_ctx *ctx, int lanes, int katoms, int ksplit);

/* Arithmetic of the mixed form's matrix-core k-space sums: MDFF_KS_F32
 * (FP32 operands) or MDFF_KS_F16X3 (split FP16 with power-of-two row
 * scaling, three products, about 22 bits).  eval_mode applies to the
 * static field and the force call, inner_mode to the low-precision SCF
 * steps; 0 keeps a setting, other values are rejected.  Defaults:
 * MDFF_KS_F16X3 for both.  No counterpart in the reference. */
enum { MDFF_KS_F32 = 2, MDFF_KS_F16X3 = 3 };
int mdff_set_kspace_modes(mdff_ctx *ctx, int eval_mode, int inner_mode);

/* Anchored-CG SCF (solver 2): the anchor, the one field evaluation at the
 * predicted dipoles that the CG steps carry along, in FP64 (anchor_mode 0)
 * or over the FP32 tables in a matrix-core mode MDFF_KS_* (then the SCF
 * makes no FP64 field evaluation; the counters book it as a low-precision
 * one).  Default: MDFF_KS_F16X3.  Split-FP16 steps take the weight scales from the CG kernels that
 * write the search direction.  No counterpart in the reference. */
int mdff_set_scf_anchor(mdff_ctx *ctx, int anchor_mode);

/* Shape of the matrix-core k-space sums over stored operands: column tiles
 * of 16 points per wavefront in the forward sum for the FP64 mode (cw64) and
 * the FP32 / split modes (cw32), each 1, 2 or 4, and the target block counts
 * of the forward and back-projection sums (their K ranges are split to reach
 * about that many blocks); 0 keeps a value.  Defaults: column tiles 1
 * (FP64), 2 (others); block targets timed on the first launch of each sum
 * shape (1/2 to 3 blocks per compute unit, the fastest kept).  No
 * counterpart in the reference (tuning). */
int mdff_set_kspace_shape(mdff_ctx *ctx, int cw64, int cw32, int fwd_blocks, int back_blocks);

/* Neighbour-list build: method 0 (default) bins the ions into a cell grid
 * and scans the cells around each ion when the box holds at least 3 cells
 * of the list cutoff per axis and 4 times as many cells as the stencil
 * around an ion, else scans all pairs; method 1 always scans all pairs (the
 * reference's gen_pbc_verletlist verlet.c:54-122 does); method 2 uses the
 * grid wherever one fits.  Rows come out in ascending neighbour index.
 * Both give the same pair set; a change applies from the next rebuild (the
 * list in place stays valid).  mdff_list_info:
 * out[0] grid in use, out[1..3] cells per axis, out[4..6] stencil
 * half-widths, out[7] row capacity.  No counterpart in the reference. */
int mdff_set_list_method(mdff_ctx *ctx, int method);
int mdff_list_info(mdff_ctx *ctx, int out[8]);

/* Mixed form: every `every`-th force call (0 never, the default), after the
 * SCF, the field of the final dipoles is formed again entirely in FP64
 * (tables, pair tensors, matrix-core sums) and the rmsd of get_rmsd_mu
 * (pim.c:431-449) evaluated on it: mdff_thermo.rmsd_fp64.  The SCF itself
 * stops on the residual of its carried low-precision field (rmsd).  No
 * counterpart in the reference (measurement support). */
int mdff_set_scf_check(mdff_ctx *ctx, int every);

/* Mixed form with the anchored CG solver: on = 1 adds an FP64 refinement
 * pass after the SCF converges on its carried residual.  The static field
 * and the field of the converged dipoles are formed in FP64, the loop
 * condition of momentpolaSCF (pim.c:236) is applied to that FP64 residual,
 * and while it is above the tolerance low-precision conjugate-gradient steps
 * carry the correction from it.  The thermo's n_scf counts both passes.
 * Off (0) by default.  No counterpart in the reference (the reference's SCF
 * is FP64 throughout). */
int mdff_set_scf_refine(mdff_ctx *ctx, int on);

/* Device timing of the last engforce by phase, in ms (HIP events on the
 * context stream): [0] short range, [1] static field, [2] SCF loop,
 * [3] final Ewald force, [4] neighbour list, [5] total. */
int mdff_get_timing(mdff_ctx *ctx, double t[6]);

/* Time one kernel of the engine at the current state: `iters` launches
 * back to back on the context stream; ms = average per launch.  The
 * matrix-core S(k) GEMM kernels (which 0, 1, 3, 4 on the matrix-core path)
 * are timed alone: one pass of the sequence with the GEMM launched `iters`
 * times back to back between a pair of HIP events (operand builds and the
 * partial-sum pass outside them); the others between two events around
 * `iters` passes.  lo = 0 the FP64 variant, 1 FP32, 3 split-FP16
 * (1 and 3 in mixed contexts only).  which:
 *   0 S(k) forward, SCF form (dipoles of the polarisable ions, 3 channels)
 *   1 back-projection, SCF form (field at the polarisable ions)
 *   2 real-space induced dipole field
 *   3 S(k) forward, force form (charges + dipoles, all ions)
 *   4 back-projection, force form (field + field gradient, all ions)
 *   5 short-range pair kernel
 *   6 real-space Ewald force kernel
 *   7 a forced rebuild of the neighbour list at the current positions by
 *     the method in use (cell grid or all pairs, lo ignored)
 * No counterpart in the reference (measurement support). */
int mdff_time_kernel(mdff_ctx *ctx, int which, int lo, int iters, double *ms);

/* Counters since creation: [0] force calls, [1] FP64 SCF field
 * evaluations, [2] inner field evaluations: the mixed form's low-precision
 * ones (the CG steps, and the anchor when it is low precision) and scfkO's
 * inner real-space sweeps, [3] polarisable ions.  No counterpart in the
 * reference (measurement support). */
int mdff_get_counters(mdff_ctx *ctx, long long out[4]);

/* The SCF solver of the PIM dipoles (algo_moment_from_pola scf).
 *   0  the reference's Jacobi schedule (momentpolaSCF pim.c:177-306) with
 *      FP64 fields: the FP64 form's default
 *   2  anchored conjugate gradients, the mixed form's default: one field
 *      evaluation of the ASPC predictor (the anchor; FP64, or a matrix-core
 *      mode set by mdff_set_scf_anchor), then CG steps on
 *      (pol^-T - T) mu = E_stat preconditioned by pol^T that apply T (low
 *      precision in the mixed form, FP64 in the FP64 form) to the search
 *      direction only and carry the field along; the rmsd of get_rmsd_mu
 *      (pim.c:431-449) is formed from the updated dipoles and field after
 *      every step, one SCF iteration per step (same fixed point as 0, fewer
 *      iterations, counts no longer the reference's)
 * Solver 2 needs every polarisability symmetric positive definite;
 * otherwise solver 0 runs.  `inner` (0-64) is the number of CG steps
 * launched before a context's first convergence poll (later SCFs launch as
 * many as the last one took).  Other solver values are rejected.  No
 * counterpart in the reference. */
int mdff_set_scf_solver(mdff_ctx *ctx, int solver, int inner);

#ifdef __cplusplus
}
#endif
#endif /* MDFF_GPU_H */
