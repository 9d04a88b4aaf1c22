// dev.h -- device-side data structures and helpers of the MI355X engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mdff_types.h"

#define NT MDFF_NTYPEMAX
#define NBR_IDX 0x3FFFFFFFu
#define NBR_SR (1u << 30)   // entry is in the short-range (non-bonded) list
#define NBR_LR (1u << 31)   // entry is in the Coulomb list
#define MDFF_PI 3.1415926535897932385
#define MDFF_TPI 6.2831853071795864770

// lattice in the form the kernels use (cell.h:6-20)
struct CellD {
    double A[3][3];  // rows a,b,c
    double B[3][3];  // reciprocal / Omega, kardir uses B[i][*].r
};

// one configuration of a batch: its cell, Ewald and cutoff settings and its
// ion and k-vector ranges in the concatenated arrays
struct BConfD {
    CellD c;
    double alpha, rc2s, rc2l;             // Ewald alpha, short and Coulomb cutoffs squared
    double selfa, selfa2, ttpiV, tfpiV;   // alpha/sqrt(pi), 2 selfa alpha^2/3, 4 pi/V, 8 pi/V
    double iv;                            // 1 / (Omega press_unit)
    int off, n, npol;                     // ions [off, off + n), polarisable count
    int Kx, Ky, Kz, koff, nk;             // k box and half-space k-vectors [koff, koff + nk)
    int boxoff;                           // offset of the full (nx >= 0) k box in the box-ordered arrays
    int awoff, wmoff;                     // offsets of the matrix-core sums' operand blocks
};

// per-type and per-type-pair tables, device resident
struct TypeTab {
    int ntype;
    int lbh, lbhd, llj, trunc;
    // BHMFTD (bhmftd.h:6-11)
    double A[NT][NT], B[NT][NT], C[NT][NT], D[NT][NT], BD[NT][NT];
    // NMLJ derived (nmlj.c:196-232)
    double sigsq[NT][NT], epsqp[NT][NT], fc[NT][NT], ptwo[NT][NT], qtwo[NT][NT], uc[NT][NT];
    double plj[NT][NT], qlj[NT][NT];
    // Tang-Toennies charge-dipole damping, ordered pair (ti,tj) = field at an
    // ion of type ti from type tj: pol_damp_*[ti][ti][tj] after symmetrisation
    int ldamp[NT][NT];  // lpoldamping[ti][ti][tj] || lpoldamping[tj][ti][tj]
    double db[NT][NT], dc[NT][NT];
    int dk[NT][NT];
    double E_TT[MDFF_MAX_TT_ORDER + 1];
    // per type
    double q[NT], mass[NT], pol[NT][9], invpol[NT][9];
    int polar[NT];
};

// scalar slots kept on the device (one double each)
enum {
    S_U_BH = 0, S_W_BH = 1,        // 6 slots W
    S_U_LJ = 7, S_W_LJ = 8,        // 6 slots
    S_U_DIR = 14, S_W_DIR = 15,    // 6 slots
    S_U_REC = 21, S_W_REC = 22,    // 6 slots
    S_QSQ = 28, S_MUSQ = 29,
    S_KIN = 30, S_UPOL = 31, S_RMSD2 = 32,
    S_ENVT = 33, S_NHC_S = 34, S_RMSD_CHK = 35,
    S_VXI = 40,                    // 16 slots
    S_XI = 56,                     // 16 slots
    S_LAMBDA = 72, S_VMEAN = 73,   // 3 slots
    S_TMP = 80,                    // 16 scratch slots
    S_COUNT = 128
};

// integer control slots
enum {
    C_REBUILD = 0, C_NUPD = 1, C_OVERFLOW = 2, C_MAXCNT = 3,
    C_SCF_ACTIVE = 4, C_ISCF = 5, C_NAN = 6, C_FIRST = 7,
    C_KO_ACTIVE = 8, C_KO_ITER = 9, C_KO_TOTAL = 10,
    C_COUNT = 16
};

__device__ __forceinline__ void min_image(const CellD &c, double sx, double sy, double sz, double &x,
                                          double &y, double &z)
{
    sx -= rint(sx);
    sy -= rint(sy);
    sz -= rint(sz);
    x = sx * c.A[0][0] + sy * c.A[1][0] + sz * c.A[2][0];
    y = sx * c.A[0][1] + sy * c.A[1][1] + sz * c.A[2][1];
    z = sx * c.A[0][2] + sy * c.A[1][2] + sz * c.A[2][2];
}

// TT_damping_functions tt_damp.c:16-32
template <typename real>
__device__ __forceinline__ void tt_damp(const double *E, real b, real c, real r, real &f, real &fd, int order)
{
    real br = b * r;
    real ex = exp(-br) * c;
    real ff = (real)E[order];
    for (int k = order - 1; k > -1; k--) ff = ff * br + (real)E[k];
    f = (real)1.0 - ff * ex;
    real p = (real)1.0;
    for (int k = 0; k < order; k++) p *= br;
    fd = (real)E[order] * p * ex * b;
}

template <int LANES>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int o = LANES / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float fast_erfc(float x) { return erfcf(x); }
__device__ __forceinline__ double fast_erfc(double x) { return erfc(x); }
