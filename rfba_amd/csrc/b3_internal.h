// b3_internal.h -- shared host/device structures of librfba_3b.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../include/rfba_3b.h"

namespace rfba3b {

// Host helpers shared by the library's translation units (b3_api.cpp):
// error message of rfba3b_last_error, device check (gfx950, CU count),
// model validation and the library-wide call lock.
int host_fail(int code, const char *fmt, ...);
int host_dev_init(int *ncu);
int host_check_model(const rfba3b_model *m);
std::mutex &host_mutex();

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
    int plane;        // chunk-local plane index (row of the per-frame sum of squares)
    int pad;
};

// One residual plane (block, q, s) of a chunk.
struct PlaneRef {
    int block, q, s, pad;
    long long rslot;
};

// Correlation task: up to kCorrSpots spots of one plane.
struct CorrTask {
    int plane;  // chunk-local plane index
    int k0;     // first spot
    int nk;     // number of spots (<= kCorrSpots)
    int item0;  // chunk-local item index of spot k0 of this plane
};

constexpr int kCorrSpots = 8;   // spots per correlation workgroup
constexpr int kCorrPix = 32;    // pixels per LDS chunk
constexpr int kCorrRow = 36;    // LDS row stride of the residual chunk (floats)
constexpr int kCorrFrames = 128;  // frames per pass (4 per thread)

struct CorrParams {
    const DevBlock *blocks;
    const PlaneRef *planes;
    const CorrTask *tasks;
    int ntasks;
    int naxmax;             // max X + Y + Z over blocks
    int Tmax;               // row stride of S0
    int NI;                 // items in the chunk (em row stride)
    const float *R;
    const double *S0;       // [plane][Tmax]: sum over pixels of R^2
    const DevModel *model;
    double *em;             // [t][NI][8]: Soff, D, diff[5]
    // Pixel-split form (nsplit > 1): virtual task v is task v / nsplit over
    // pixels [split pstep, (split + 1) pstep) of its plane, split = v %
    // nsplit; instead of em it writes its raw sums, which b3_corr_reduce adds
    // in a fixed order and turns into em.
    int nsplit;             // 0 or 1: whole planes
    int pstep;              // pixels per split (a multiple of kCorrPix)
    double *part;           // [nsplit][Tmax][NI][6]: C, G_0..4
    double *part_qh;        // [nsplit][NI][6]: Q, H_0..4
};



struct Item {
    int block, q, k, s;
    long long rslot;
};

// b3_corr_reduce: items of one block with Q = 1 whose plane index is the
// item's sample s (the resident functor's layout).
struct CorrReduceParams {
    const DevBlock *block;
    const Item *items;
    int NI, T, nsplit, Tmax;
    const double *part, *part_qh, *S0;
    double *em;
};

struct HmmParams {
    const DevBlock *blocks;
    const Item *items;      // chunk's items (chunk-local index 0..NI)
    int lo, n;              // items lo .. lo + n - 1 of the chunk
    int NI;
    int ckpt;               // 0: ceil(sqrt(T)) per block
    int ckslots;            // doubles per item in ck: 3 ceil(T/ckpt) + 2 ckpt
    const double *em;
    double *ck;             // [slot][NI]
    const DevModel *model;
    double *delta;          // b3_hmm_wave MODE_DELTA: [chunk-local item][T][3]
};

// b3_hmm_wave: one wave per item, up to kWaveFrames frames per lane.
constexpr int kWaveFrames = 8;
constexpr int kWaveTmax = 64 * kWaveFrames;

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

// FFBS Hessian (sampled_background_spot_hessian_ffbs): per-pixel table of
// the spot's value, the 15 upper entries of its Hessian ((0,0), (0,1), ...,
// (0,4), (1,1), ..., (4,4)) and its 5 derivatives, FP32, padded to 24.
constexpr int kHessBasis = 24;
constexpr int kHessUpper = 15;
constexpr int kHessParts = 20;   // per frame: sum e H (15), sum e d (5)
constexpr int kHessFrames = 4;   // frames per parts workgroup

struct HessParams {
    const DevBlock *block;  // region box and mask; params[0] = the spot
    const DevModel *model;
    float *basis;           // [box pixels][kHessBasis]
    double *fixed;          // [basis workgroups][15]: sums of d_i d_j (i <= j)
    const float *R;         // residual planes [S][T][box pixels]
    int S, T;
    double *parts;          // [S][T][kHessParts]
};

// ---------------------------------------------------------------------------
// Extended-range values: v = m 2^e with m in [0.5, 1) and an exponent of
// its own, an integer held in a double (exact to 2^53; -inf for m = 0).
// Forward messages and transfer-matrix products are kept this way so that
// no weight underflows: a state whose weight is e^-10000 relative to
// another keeps it, as in the reference's log-space forward algorithm
// (forward_algorithm.h), and sums drop a term only when it is below
// 2^-1074 of the largest, as logsumexp's exp() does.
__device__ __forceinline__ void xsplit(double v, double e0, double &m, double &e) {
    int k;
    m = frexp(v, &k);
    e = v > 0.0 ? e0 + (double)k : -__builtin_inf();
}

// 2^(e - E) m for e <= E (E the largest exponent; -inf - -inf gives 0 for a
// zero m)
__device__ __forceinline__ double xalign(double m, double e, double E) {
    return ldexp(m, (int)fmin(fmax(e - E, -2000.0), 0.0));
}

// m0 2^e0 + m1 2^e1 + m2 2^e2, the terms aligned to the largest exponent of
// a nonzero term (a term m_i = 0 may carry any exponent, e.g. after a
// product with a zero transition probability)
__device__ __forceinline__ void xsum3(double m0, double e0, double m1, double e1, double m2, double e2, double &m,
                                      double &e) {
    const double ninf = -__builtin_inf();
    const double E = fmax(m0 > 0.0 ? e0 : ninf, fmax(m1 > 0.0 ? e1 : ninf, m2 > 0.0 ? e2 : ninf));
    double s = xalign(m0, e0, E);
    s += xalign(m1, e1, E);
    s += xalign(m2, e2, E);
    xsplit(s, E, m, e);
}

// xalign for e <= E (E the largest of the exponents, zero terms included):
// e - E is at most 0, or NaN for -inf - -inf (a zero term), which fmax
// turns into -2000.
__device__ __forceinline__ double xalign_le(double m, double e, double E) {
    return ldexp(m, (int)fmax(e - E, -2000.0));
}

// xsum3 for products of values that xsplit formed, where a zero term
// always carries exponent -inf (no masking needed)
__device__ __forceinline__ void xsum3n(double m0, double e0, double m1, double e1, double m2, double e2, double &m,
                                       double &e) {
    const double E = fmax(e0, fmax(e1, e2));
    double s = xalign_le(m0, e0, E);
    s += xalign_le(m1, e1, E);
    s += xalign_le(m2, e2, E);
    xsplit(s, E, m, e);
}

// e^y for y <= 0: y = n ln2 + r with |r| <= ln2 / 2 (Cody and Waite's
// reduction, ln 2 in two parts, each product with n subtracted by a fused
// multiply-add), e^r by exp().
__device__ __forceinline__ void xexp(double y, double &m, double &e) {
    const double n = rint(y * 1.4426950408889634074);
    const double r = fma(-n, 1.90821492927058770002e-10, fma(-n, 6.93147180369123816490e-01, y));
    xsplit(exp(r), n, m, e);
}

// Subtract the largest exponent of the nonzero entries from all of them.
template <int N>
__device__ __forceinline__ void xshift(double (&e)[N]) {
    double E = -__builtin_inf();
#pragma unroll
    for (int i = 0; i < N; i++) E = fmax(E, e[i]);
    if (E > -__builtin_inf()) {
#pragma unroll
        for (int i = 0; i < N; i++) e[i] -= E;
    }
}

}  // namespace rfba3b
