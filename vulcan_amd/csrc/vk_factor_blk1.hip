// Blocked Gauss-Jordan chain factorisation compiled for one resident workgroup per CU
// (factor variant 1).
#include "vk_factor_blk.h"

int vk_launch_factor_blocked1(int nb, int ni, int ncols, const int *off, const int *order, const double *D,
                              const double *Lsub, const double *Usup, double *Dinv, int *info, const vk_colstate *st,
                              hipStream_t s) {
    return launch_blocked<1>(nb, ni, ncols, off, order, D, Lsub, Usup, Dinv, info, st, s);
}
