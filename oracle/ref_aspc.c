/*
 * ref_aspc.c -- harness around the reference's own ASPC predictor
 * (extrapolate_dipole_aspc, pim.c:314-424) for the oracle pins (TEST
 * INFRASTRUCTURE).  Linked against libmdff20_ref.so; the predictor reads
 * only the globals nion, istep, extrapolate_order and dipia_ind, which are
 * set here for one ion.
 *
 * usage: ref_aspc <extrapolate_order> <istep> <h0x> <h0y> <h0z> ... <h5z>
 * prints the predicted dipole and the history after the predictor's shift,
 * 21 numbers at %.17g on one line.
 */
#include <stdio.h>
#include <stdlib.h>

#include "config.h"
#include "io.h"
#include "md.h"
#include "pim.h"

int main(int argc, char **argv)
{
    const int nh = MAX_EXTRAPOLATE_ORDER + 1;
    if (argc != 3 + 3 * nh) {
        fprintf(stderr, "usage: %s order istep 18 history values\n", argv[0]);
        return 2;
    }
    ionode = 1;
    nion = 1;
    extrapolate_order = atoi(argv[1]);
    istep = atoi(argv[2]);
    dipia_ind = malloc(sizeof(*dipia_ind));
    for (int k = 0; k < nh; k++)
        for (int i = 0; i < 3; i++) dipia_ind[0][k][i] = atof(argv[3 + 3 * k + i]);
    double mu[1][3] = {{0.0, 0.0, 0.0}};
    extrapolate_dipole_aspc(mu, NULL, 1);
    for (int i = 0; i < 3; i++) printf("%.17g ", mu[0][i]);
    for (int k = 0; k < nh; k++)
        for (int i = 0; i < 3; i++) printf("%.17g ", dipia_ind[0][k][i]);
    printf("\n");
    return 0;
}
