// b3_internal.h -- shared host/device structures of librfba_3b.so.
#pragma once
#include <stdint.h>

#include "../../include/rfba_3b.h"

namespace rfba3b {

enum { MODE_LOGL = 0, MODE_GRAD = 1, MODE_DELTA = 2 };

struct DevModel {
    double A[9];
    double pi[3];
    double variance;
    double zfactor;
    int z_exact;
    int pad;
};

// One block on the device: a dense X x Y x Z box whose local voxel (0,0,0)
// has coordinates (ox, oy, oz).
struct DevBlock {
    int X, Y, Z, T, K, Q, S;
    int ox, oy, oz;
    int npix;               // pixels in the region (P, or the mask's count)
    double b0;
    const float *data;      // [T][P]
    const double *params;   // [Q][K][5]
    const uint8_t *states;  // [S][K][T]
    const uint8_t *mask;    // [P] or null (1 = pixel belongs to the region)
    double *logL;           // [Q][K][S]
    double *grad;           // [Q][K][S][5]
};

struct PlaneTask {
    int block, q, s, t;
    long long rslot;  // residual plane slot
};

struct Item {
    int block, q, k, s;
    long long rslot;
};

struct EvalParams {
    const DevBlock *blocks;
    const Item *items;
    int nitems;
    int *counter;
    const float *R;         // residual planes, slot stride T*P of the item's block
    const DevModel *model;
    int rmode;              // 0: R excludes nothing (add the spot's own sampled term); 1: R = data - bg
    int ckpt;
    int Tmax, XYmax, Zmax;
    double *delta_out;      // MODE_DELTA: [item][T][3]
};

}  // namespace rfba3b
