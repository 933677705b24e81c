// srrf_internal.h -- shared host/device structures of librfba_srrf.so.
#pragma once
#include <stdint.h>

#include "../../include/rfba_srrf.h"

namespace rfba {

// Sampling tables.  The sample list is kept in its given order; samples are
// grouped into
//   runs:   maximal consecutive stretches with the same (x, y) offset
//   groups: maximal consecutive stretches of runs with the same x offset
// so that x-interpolated planes (per group) and y-interpolated columns (per
// run) can be shared while every voxel still sums its samples in list order.
// Two consecutive groups whose x offsets a and b = a + 1 are multiples of
// 1/2 are merged: the sample x coordinate (X + 1/2) + b of SR column X is
// exactly ((X + 1) + 1/2) + a, so the runs of b read the planes of a one
// column over, and the merged group's x pass covers kTX + 1 columns.
constexpr int kMaxSamples = RFBA_SRRF_MAX_SAMPLES;
constexpr int kMaxOffsets = 256;  // distinct offsets per axis

struct DevSamples {
    int n;
    float divisor;
    float R;        // ring radius (SR pixels)
    float invR;     // 1/R when R is a power of two (then x/R == x*invR)
    int r_pow2;
    int fast_ok;    // R in [2^-20, 2^20]: the kernel's branch-free path applies
    int nox, noy, noz;
    int nruns, ngroups;
    float oxmin, oxmax, oymin, oymax, ozmin, ozmax;  // include 0 (voxel centre)
    float ox[kMaxOffsets], oy[kMaxOffsets], oz[kMaxOffsets];
    // per sample (list order)
    float rx[kMaxSamples], ry[kMaxSamples], rz[kMaxSamples];
    int siz[kMaxSamples];
    // runs (shift 1, with merged groups: reads the group's x-interpolated
    // planes one column over)
    int run_start[kMaxSamples], run_len[kMaxSamples], run_iy[kMaxSamples], run_shift[kMaxSamples];
    // groups
    int grp_ix[kMaxSamples], grp_run0[kMaxSamples], grp_nrun[kMaxSamples];
    // the groups with merges (ext 1: x pass over kTX + 1 columns); a launch
    // uses them when LaunchParams::merge is set (the planner's choice: only
    // where the wider planes cost no occupancy)
    int mgrp_ix[kMaxSamples], mgrp_run0[kMaxSamples], mgrp_nrun[kMaxSamples], mgrp_ext[kMaxSamples];
    int nmgroups, any_ext;
};

// One stack (or block of a source series) and the SR box of it to compute.
struct DevItem {
    const void *src;     // source series base (element type per launch)
    long long sy, sz, st;  // source strides in elements (x stride 1)
    int ox, oy, oz;      // stack origin inside the source (low-res)
    int W, H, D, T;      // stack dims (low-res) and frame count
    int M;
    int pad0;
    float *out;          // SR (0,0,0) of this stack
    long long osy, osz;  // output strides in floats
    int bx0, by0, bz0, bx1, by1, bz1;  // SR box [b0, b1) to write
};

struct DevTile {
    int item;
    int X0, Y0, Z0;  // SR tile origin (stack-local)
};

// Compile-time tile shape: 16 x 16 columns per workgroup (one thread per
// column), TZ voxels along z per thread.
constexpr int kTX = 16;
constexpr int kTY = 16;
constexpr int kThreads = kTX * kTY;

struct LaunchGeom {
    int NU, NV, NW;      // gradient-plane extents held in LDS
    int hx, hy, hz;      // gradient halo
    int RU, RV, RW;      // raw extents = N + 2h
    int TXP;             // padded X stride of the x-interpolated planes (row layout)
    // Column layout of the x-interpolated planes (pxv = 1): Px[c][v][X][w],
    // the w values of a column contiguous (NWP = NW rounded up to even) and
    // planes of one v VS floats apart (VS = 16 NWP + 16: rows of v shifted by
    // 16 banks; column 16 of a merged group lies in those 16 when NWP <= 16,
    // else VS = 17 NWP + 16), so a thread's y pass reads pairs of w from one
    // base.
    int pxv, NWP, VS;
    int noz;
    // float offsets into dynamic LDS
    int off_raw, off_g, off_px, off_py, off_xtab, off_ztab, off_zc, off_ctab, off_dsum, off_stab, off_total;
    // sample tables inside the off_stab region (words): ox at 0
    int st_oy, st_oz, st_grp, st_run, st_smp;
};

struct LaunchParams {
    const DevSamples *smp;
    const DevItem *items;
    const DevTile *tiles;
    int ntiles;
    int *counter;        // work-queue head (zeroed by a kernel before launch)
    int *err;            // geometry-violation flag
    float *scratch;      // per-workgroup frame series (TRAC / TRPPM)
    long long scratch_stride;  // floats per workgroup slot
    int Tmax;
    // configuration shared by all items
    int method, itc, order, remove_con, renorm, smooth, iw;
    int mR, xR, yR, zR;
    // smoothed gradient of integer input from shared box sums: every sum has
    // at most 256 terms and the box sums fit in the x-plane area
    int sep;
    // the merged sample groups (DevSamples::mgrp_*) and run shifts apply
    int merge;
    LaunchGeom g;
};

}  // namespace rfba
