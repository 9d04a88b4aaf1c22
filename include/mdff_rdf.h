/*
 * mdff_rdf.h -- C ABI of the device radial distribution histogram and the
 * device trajectory of libmdff_gpu.so.  Plain C types only.
 *
 * Reference interfaces replaced (FilipeVasconcelos/mdff20, src/):
 *   mdff_rdf_add_frame / mdff_rdf_sample / mdff_rdf_sample_trajectory
 *       <- void get_gr(void)   radial_distribution.h:22, radial_distribution.c:180-231
 *          (the per-frame histogram rdf_readtraj runs, radial_distribution.c:75-99)
 *   mdff_rdf_gr
 *       <- the normalisation of int write_rdf(void)   radial_distribution.c:120-141
 *   mdff_set_trajectory / mdff_traj_frame
 *       <- the TRAJFF records of run_md (md.c:275-277): the frames stay in
 *          device memory during the run and are read back when wanted
 * The same-name drop-in get_gr() over the reference's globals is in
 * mdff_dropin.h.
 */
#ifndef MDFF_RDF_H
#define MDFF_RDF_H

#include "mdff_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdff_rdf mdff_rdf;

/* default_rdf (radial_distribution.c:37-42) for cell A (rows a, b, c):
 * cutgr = half the smallest width of lattice (cell.c:43-47), nbins =
 * (int)(cutgr / 0.1) + 1 (resg 0.1) */
int mdff_rdf_defaults(const double A[3][3], double *cutgr, int *nbins);

/* A histogram of n ions of ntype types (type[i] in [0, ntype)): nbins bins
 * of width resg, pairs closer than cutgr (get_gr radial_distribution.c:190-
 * 216; default_rdf :37-42 gives cutgr = half the smallest cell width and
 * nbins = (int)(cutgr / 0.1) + 1).  nbins must hold every distance below
 * cutgr (nbins >= floor(cutgr / resg) + 1), else NULL with MDFF_ERR_ARG in
 * *status.  The handle owns its device memory and stream on the current
 * device. */
mdff_rdf *mdff_rdf_create(int n, int ntype, const int *type, int nbins, double resg, double cutgr, int *status);
void mdff_rdf_destroy(mdff_rdf *h);
/* zero the counts and the frame count */
int mdff_rdf_reset(mdff_rdf *h);

/* One frame from the host (a TRAJFF record, read_next_traj config.c:448-
 * 502): positions rx, ry, rz of cell A (rows a, b, c).  With write_back the
 * positions are replaced by what get_gr leaves in the globals: kardir
 * (cell.c:125-143) then dirkar (cell.c:146-161), formed on the device. */
int mdff_rdf_add_frame(mdff_rdf *h, const double A[3][3], double *rx, double *ry, double *rz, int write_back);
/* One frame of an engine's current positions, read in device memory (the
 * engine's types must be the handle's). */
int mdff_rdf_sample(mdff_rdf *h, mdff_ctx *ctx);
/* Every frame of an engine's device trajectory (mdff_set_trajectory), in
 * one launch. */
int mdff_rdf_sample_trajectory(mdff_rdf *h, mdff_ctx *ctx);

/* Measurement: the device time of one frame of the engine's positions
 * (kardir and pair kernels), averaged over iters frames timed with HIP
 * events on the engine's stream; the counts are reset afterwards.  No
 * counterpart in the reference. */
int mdff_rdf_time(mdff_rdf *h, mdff_ctx *ctx, int iters, double *ms);

/* Ordered pair counts counts[nbins][ntype][ntype] of all frames so far, as
 * gr[igr][it1][it2] of get_gr (radial_distribution.c:220), without its
 * extra all-pair increment of gr[igr][0][0] (:219, which the drop-in get_gr
 * adds).  info (may be NULL): [0] frames, [1] pairs whose bin fell outside
 * [0, nbins) (not counted). */
int mdff_rdf_counts(mdff_rdf *h, long long *counts, long long info[2]);
/* write_rdf's normalisation (radial_distribution.c:120-141) of ordered
 * counts over nconf frames in a cell of volume omega: r[igr] (may be NULL)
 * and g[igr][it1][it2] = counts / (nconf vol n_it1 n_it2), with vol the
 * shell volume 4 pi (resg r^2 + resg^3 / 12) / omega at the bin centre and
 * n_it the handle's ion counts per type. */
int mdff_rdf_gr(const mdff_rdf *h, const long long *counts, long long nconf, double omega, double *r, double *g);

/* Device trajectory of an engine.  With tprint > 0, mdff_md_step and
 * mdff_md_run copy r, v and f into device memory after every step that
 * run_md writes to TRAJFF (iopsnode(istep, npas, tprint, tstart), io.c:67-
 * 69: istep % tprint == 0 or istep == npas, and istep >= tstart), up to
 * `capacity` frames (9 n doubles each); later frames are counted as dropped.
 * tprint = 0 turns it off and frees the frames.  No host copy during the
 * steps. */
int mdff_set_trajectory(mdff_ctx *ctx, int tprint, int tstart, int npas, int capacity);
/* out[0] frames held, out[1] capacity, out[2] frames dropped */
int mdff_traj_info(mdff_ctx *ctx, int out[3]);
/* frame k: its step and, for each non-NULL pointer, the x, y, z arrays of
 * positions r[3], velocities v[3] and forces f[3] */
int mdff_traj_frame(mdff_ctx *ctx, int k, int *istep, double *r[3], double *v[3], double *f[3]);
/* forget the frames held (keeps the buffer) */
int mdff_traj_clear(mdff_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif
