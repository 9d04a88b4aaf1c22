/*
 * mdff_types.h -- plain-data types shared by the host C front end, the HIP
 * engine (libmdff_gpu.so) and the CPU oracle.
 *
 * mdff_params holds every control-file keyword of mdff20's force and
 * integration path, in control-file units (Angstrom, eV, amu, e, ps, K).
 * Field meanings and defaults follow the reference parsers:
 *   global.c:13-94   (Fposff, lverletL, skindiff, cutoffs, lstatic, lreduced)
 *   md.c:29-118      (integrator, dt, temp, npas, print periods, NHC)
 *   field.c:31-183   (lnmlj/lbhmft/lbhmftd/lcoulombic, per-type tables,
 *                     damping tables, Ewald alpha/kmax/epsw/lautoES)
 *   bhmftd.c:26-90, nmlj.c:39-107, nonbonded.c:7-28, pim.c:28-85
 * Per-type-pair tables are stored full (both triangles) after the same
 * symmetrisation the reference applies (bhmftd.c:98-110, nmlj.c:205-212,
 * field.c:209-219).
 */
#ifndef MDFF_TYPES_H
#define MDFF_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

#define MDFF_NTYPEMAX 16
#define MDFF_MAX_LEN 80
#define MDFF_MAX_EXTRAPOLATE_ORDER 5
#define MDFF_MAX_TT_ORDER 8          /* E_TT has 9 entries: tt_damp.h:4 */

/* integrator ids, md.c:102-105 */
enum { MDFF_NVE_LF = 0, MDFF_NVE_VV = 1, MDFF_NVT_NHCN = 2 };
/* NMLJ truncation, nmlj.c:213-215 */
enum { MDFF_TRUNC_NO = 0, MDFF_TRUNC_LINEAR = 1, MDFF_TRUNC_QUADRATIC = 2 };
/* PIM algorithms, pim.c:73-74; MDFF_PIM_PCG is this engine's fast mode */
enum { MDFF_PIM_SCF = 0, MDFF_PIM_SCFKO = 1, MDFF_PIM_PCG = 2 };

/* status codes returned by every entry point (0 = ok) */
enum {
    MDFF_OK = 0,
    MDFF_ERR_ARG = 1,          /* bad argument / inconsistent sizes        */
    MDFF_ERR_CONFIG = 2,       /* invalid control-file value               */
    MDFF_ERR_IO = 3,           /* file missing / unreadable                */
    MDFF_ERR_DEVICE = 4,       /* HIP runtime error                        */
    MDFF_ERR_LIST_OVERFLOW = 5,/* neighbour list capacity exceeded         */
    MDFF_ERR_NAN = 6,          /* non-finite energy or force               */
    MDFF_ERR_SCF = 7,          /* SCF stopped at max iterations unconverged
                                  (reported, not fatal: the reference accepts
                                  it silently, pim.c:236)                  */
    MDFF_ERR_UNSUPPORTED = 8,  /* feature not built in this engine         */
    MDFF_ERR_NOMEM = 9,
    MDFF_ERR_SCRATCH = 10      /* larger than the fixed scratch of a
                                  no-allocation entry point                */
};

typedef struct mdff_params {
    int ntype;                                   /* from POSFF              */
    /* ---- global.c ---- */
    int Fposff;                                  /* 0 rnn,1 rvn,2 rvf (2)   */
    int lverletL;                                /* default 1               */
    double skindiff;                             /* default 0.15            */
    int lstatic, lrdf, lpstress, lreduced;
    double cutshortrange, cutlongrange;
    /* ---- md.c ---- */
    int integrator;                              /* default nve-vv          */
    double dt, temp, tauTberendsen, timesca_thermo, timesca_baro; /* ps, K */
    int npas, nprint, fprint, cprint, tprint, tstart, nequil, nequilT;
    int nhc_n, nhc_mults, nhc_yosh_order;
    /* ---- field.c ---- */
    int lnmlj, lbhmft, lbhmftd, lcoulombic, lsymmetric;
    int nmassit, nqit, ndipit, nquadit, npolit;   /* counts of entries read */
    double massit[MDFF_NTYPEMAX];
    double qit[MDFF_NTYPEMAX];
    double dipit[MDFF_NTYPEMAX][3];
    double quadit[MDFF_NTYPEMAX][3][3];
    double polit[MDFF_NTYPEMAX][3][3];
    int lpolar[MDFF_NTYPEMAX];
    /* damping tables [polarisable type][it][jt] (config.h:27-32) */
    int lpoldamping[MDFF_NTYPEMAX][MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    double pol_damp_b[MDFF_NTYPEMAX][MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    double pol_damp_c[MDFF_NTYPEMAX][MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    int pol_damp_k[MDFF_NTYPEMAX][MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    double alphaES;
    int kES[3];
    double epsw;
    int lautoES;
    /* ---- bhmftd.c ---- */
    double Abhmftd[MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    double Bbhmftd[MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    double Cbhmftd[MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    double Dbhmftd[MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    double BDbhmftd[MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    /* ---- nmlj.c ---- */
    double sigmalj[MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    double epslj[MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    double plj[MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    double qlj[MDFF_NTYPEMAX][MDFF_NTYPEMAX];
    int trunc;
    /* ---- pim.c ---- */
    int algo_pim, algo_extrapolate_dipole, extrapolate_order;
    int min_scf_pim_iter, max_scf_pim_iter;
    double conv_tol_ind, omegakO;
} mdff_params;

/* POSFF / CONTFF content (config.c:291-446, layout documented in
 * bin/config.py:207-241).  Arrays are owned by the struct (mdff_config_free). */
typedef struct mdff_config {
    int nion, ntype;
    char name[MDFF_MAX_LEN + 1];
    double A[3][3];                      /* rows = lattice vectors a,b,c   */
    char atypit[MDFF_NTYPEMAX][MDFF_MAX_LEN + 1];
    int nionit[MDFF_NTYPEMAX];
    int *typia;                          /* by count order, config.c:135-148 */
    double *rx, *ry, *rz, *vx, *vy, *vz, *fx, *fy, *fz;
} mdff_config;

/* observables of one force evaluation / MD step (thermo.h:3-14) */
typedef struct mdff_thermo {
    int istep;
    double e_kin, e_nvt;
    double u_lj, u_bhmftd, u_coul, u_pol;
    double pvir_lj, pvir_bhmftd, pvir_coul;
    double tau_lj[3][3], tau_bhmftd[3][3], tau_coul[3][3];
    /* derived, thermo.c:16-33 */
    double temp_r, u_tot, e_tot, h_tot, pressure;
    /* engine diagnostics */
    int n_scf;             /* SCF iterations in the last force call */
    double rmsd;           /* final SCF residual (pim.c:431-449 definition); in the
                              mixed form the residual of the carried low-precision
                              field the SCF stopped on */
    double rmsd_fp64;      /* mixed form with mdff_set_scf_check: the same residual
                              with the field of the final dipoles formed in FP64;
                              -1 when not evaluated in this call */
    int verlet_updates;    /* list rebuilds so far (verlet.c:152-154)  */
} mdff_thermo;

#ifdef __cplusplus
}
#endif
#endif /* MDFF_TYPES_H */
