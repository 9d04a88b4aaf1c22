// typetab.h -- per-type tables of the device kernels from the control
// parameters: BHMFTD (bhmftd.h:6-11), NMLJ derived constants (init_nmlj
// nmlj.c:196-232), Tang-Toennies charge-dipole damping (field.c, pim.c:28-85
// keywords), charges, masses and polarisabilities with their inverse.
#pragma once
#include <math.h>
#include <string.h>

#include "dev.h"

static inline void typetab_build(const mdff_params &p, int ntype, double srcutsq, TypeTab &T)
{
    memset(&T, 0, sizeof(T));
    T.ntype = ntype;
    T.lbh = p.lbhmft || p.lbhmftd;
    T.lbhd = p.lbhmftd;
    T.llj = p.lnmlj;
    T.trunc = p.trunc;
    double two16 = pow(2.0, 1.0 / 6.0);
    for (int i = 0; i < NT; i++)
        for (int j = 0; j < NT; j++) {
            T.A[i][j] = p.Abhmftd[i][j];
            T.B[i][j] = p.Bbhmftd[i][j];
            T.C[i][j] = p.Cbhmftd[i][j];
            T.D[i][j] = p.Dbhmftd[i][j];
            T.BD[i][j] = p.BDbhmftd[i][j];
            // init_nmlj nmlj.c:196-232
            double pp = p.plj[i][j], qq = p.qlj[i][j];
            T.plj[i][j] = pp;
            T.qlj[i][j] = qq;
            T.epsqp[i][j] = p.epslj[i][j] / (qq - pp);
            T.sigsq[i][j] = two16 * two16 * p.sigmalj[i][j] * p.sigmalj[i][j];
            T.fc[i][j] = (pp * qq * T.epsqp[i][j]) / T.sigsq[i][j];
            T.ptwo[i][j] = pp * 0.5;
            T.qtwo[i][j] = qq * 0.5;
            if (srcutsq > 0) {
                double sr2 = T.sigsq[i][j] / srcutsq, sr = sqrt(sr2);
                T.uc[i][j] = T.epsqp[i][j] * (pp * pow(sr, qq) - qq * pow(sr, pp));
            }
            T.ldamp[i][j] = (p.lpoldamping[i][i][j] || p.lpoldamping[j][i][j]) ? 1 : 0;
            T.db[i][j] = p.pol_damp_b[i][i][j];
            T.dc[i][j] = p.pol_damp_c[i][i][j];
            T.dk[i][j] = p.pol_damp_k[i][i][j];
        }
    T.E_TT[0] = 1.0;
    for (int k = 1; k <= MDFF_MAX_TT_ORDER; k++) T.E_TT[k] = T.E_TT[k - 1] / (double)k;
    for (int t = 0; t < NT; t++) {
        T.q[t] = p.lcoulombic ? p.qit[t] : 0.0;
        T.mass[t] = p.massit[t];
        T.polar[t] = (p.lcoulombic && p.lpolar[t]) ? 1 : 0;
        for (int k = 0; k < 9; k++) T.pol[t][k] = T.polar[t] ? p.polit[t][k / 3][k % 3] : 0.0;
        if (T.polar[t]) {
            // invertmatrix3x3 inverse_matrix.c:6-23, closed-form adjugate
            const double *m = T.pol[t];
            double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                         m[2] * (m[3] * m[7] - m[4] * m[6]);
            double id = 1.0 / det;
            T.invpol[t][0] = (m[4] * m[8] - m[5] * m[7]) * id;
            T.invpol[t][1] = (m[2] * m[7] - m[1] * m[8]) * id;
            T.invpol[t][2] = (m[1] * m[5] - m[2] * m[4]) * id;
            T.invpol[t][3] = (m[5] * m[6] - m[3] * m[8]) * id;
            T.invpol[t][4] = (m[0] * m[8] - m[2] * m[6]) * id;
            T.invpol[t][5] = (m[2] * m[3] - m[0] * m[5]) * id;
            T.invpol[t][6] = (m[3] * m[7] - m[4] * m[6]) * id;
            T.invpol[t][7] = (m[1] * m[6] - m[0] * m[7]) * id;
            T.invpol[t][8] = (m[0] * m[4] - m[1] * m[3]) * id;
        }
    }
}
