// srrf_kernel.hip -- fused 3D SRRF radiality kernel for gfx950 (MI355X).
//
// One persistent workgroup of 256 threads takes SR tiles (16 x 16 columns x
// TZ voxels) from a device work queue and, for every frame of the tile's
// stack, runs
//   raw halo (uint16 or float, coalesced loads, clamp-to-edge)
//     -> LDS -> xyz gradient in LDS                (gradKernel, kernel.cu:44-125)
//   -> per x-offset group: x-interpolated planes in LDS
//   -> per (x,y)-offset run: y-interpolated columns (thread-private LDS)
//   -> per sample: z-pass, radiality term         (interpolateG :141-169,
//                                                   calculateRadiality :186-253)
//   -> temporal accumulation in registers          (:291-555)
// The per-voxel signed sample sums live in thread-private LDS rows so that
// the voxel loop stays rolled (register pressure).
// The tricubic Catmull-Rom interpolation of the reference evaluates
// p = sum_i G*wx, q = sum_j p*wy, s = sum_k q*wz in exactly this order, so
// the partial sums p (depending on the sample's x only) and q (x and y only)
// are shared between samples and voxels without changing a single rounding.
// Every voxel still accumulates its sample terms in list order.  The
// translation unit is compiled with -ffp-contract=off and IEEE division and
// square root, so each voxel-frame value is the reference arithmetic.
//
// TRAC and TRPPM need the whole per-voxel series after the temporal mean is
// known (two-pass in the reference, :458-507); the series goes to a
// per-workgroup HBM slot and is reduced per thread after the frame loop.
#include <hip/hip_runtime.h>

#include "srrf_internal.h"

namespace rfba {

__device__ __forceinline__ float cubic_w(float xx) {
    // cubic, kernel.cu:128-138 (Catmull-Rom, a = 0.5)
    const float a = 0.5f;
    float x = xx;
    if (x < 0.0f) x = -x;
    float z = 0.0f;
    if (x < 1.0f)
        z = x * x * (x * (-a + 2.0f) + (a - 3.0f)) + 1.0f;
    else if (x < 2.0f)
        z = -a * x * x * x + 5.0f * a * x * x - 8.0f * a * x + 4.0f * a;
    return z;
}

// Sample coordinate c (SR pixels) -> first tap index b = u0 - 1, the four
// Catmull-Rom weights and the in-range test of interpolateG (:143-151).
__device__ __forceinline__ void axis_weights(float c, int M, int dim, int &b, float &w0,
                                             float &w1, float &w2, float &w3, int &valid) {
    float x = c / (float)M;
    valid = !(x < 1.5f || x > (float)dim - 1.5f);
    int u0 = (int)floorf(x - 0.5f);
    w0 = cubic_w(x - ((float)(u0 - 1) + 0.5f));
    w1 = cubic_w(x - ((float)(u0) + 0.5f));
    w2 = cubic_w(x - ((float)(u0 + 1) + 0.5f));
    w3 = cubic_w(x - ((float)(u0 + 2) + 0.5f));
    b = u0 - 1;
}

__device__ __forceinline__ int first_tap(float c, int M) {
    return (int)floorf(c / (float)M - 0.5f) - 1;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// e / d for 0 <= e < 2^16 and d >= 1, with rd the correctly rounded 1 / d:
// (e + 0.5) / d lies at least 0.5 / d from an integer, and the two roundings
// of (e + 0.5) rd move it by less than 2^-23 (e + 0.5) / d < 0.01 / d.  The
// element loops of a frame use it for their index decompositions (LDS
// extents: e < 40960), where integer division by a runtime divisor costs
// about 25 instructions.
__device__ __forceinline__ int qdiv(int e, float rd) { return (int)(((float)e + 0.5f) * rd); }

// uint16 input: raw values are integers below 2^16
template <typename Tin> constexpr bool kIntInput = false;
template <> constexpr bool kIntInput<uint16_t> = true;

// Signed radiality term of one sample (kernel.cu:227-240 with _calculateDk
// :174-183): -Dk when diverging, +Dk otherwise (dd - Dk == dd + (-Dk)).
struct TermIn {
    float rx, ry, rz;     // ring offset of the sample
    float dx, dy, dz;     // x0 - Xc, y0 - Yc, z0 - Zc
    float x0, y0, z0;     // sample position
    float Xc, Yc, Zc;     // voxel centre
};

// General form: IEEE sqrtf and division for every input.
template <bool RP2>
__device__ __forceinline__ float term_general(float Gx, float Gy, float Gz, const TermIn &a, float R, float invR) {
    float g2 = Gx * Gx + Gy * Gy;
    g2 = g2 + Gz * Gz;
    const float GMag = sqrtf(g2);
    float num = Gx * a.rx + Gy * a.ry;
    num = num + Gz * a.rz;
    const float q = num / (GMag * R);
    const bool dv = (q > 0.0f) || (q != q);
    float tn = Gx * a.dx + Gy * a.dy;
    tn = tn + Gz * a.dz;
    const float tt = tn / (GMag * GMag);
    const float ex = a.x0 - (Gx * tt + a.Xc), ey = a.y0 - (Gy * tt + a.Yc), ez = a.z0 - (Gz * tt + a.Zc);
    float d2 = ex * ex + ey * ey;
    d2 = d2 + ez * ez;
    float Dk = sqrtf(d2);
    Dk = 1.0f - (RP2 ? Dk * invR : Dk / R);
    Dk = Dk * Dk;
    return dv ? -Dk : Dk;
}

// Branch-free form, bit-identical to term_general whenever `bad` stays
// false: both sqrt arguments lie in [2^-96, inf) (there the compiler's sqrtf
// reduces to v_sqrt_f32 plus the FMA-residual choice between neighbours) and
// the sign of num / (GMag * R) is decided without dividing.  With g2 in
// [2^-96, 2^100] and R in [2^-20, 2^20] (checked on the host, else every
// sample takes the general form), den = GMag * R lies in [2^-68, 2^70]; a
// numerator >= 2^-60 then gives a quotient >= 2^-130, which rounds to a
// positive value, and a numerator <= 0 a quotient that is not > 0.  A term
// not flagged has finite G (NaN or infinite G makes g2 NaN or infinite), so
// a finite numerator: "diverging" is num > 0 unless 0 < num < 2^-60.
// G2B: g2 <= 2^100 holds for every input (integer input: each gradient
// component is at most 2 * 8 * 17 * 17 terms of at most 65535, times
// tricubic weights whose absolute sums are below 2.2, so g2 < 2^61), and
// the upper test is left out.
__device__ __forceinline__ float sqrt_core(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __int_as_float(__float_as_int(s) - 1);
    const float su = __int_as_float(__float_as_int(s) + 1);
    float r = (fmaf(-sd, s, x) <= 0.0f) ? sd : s;
    r = (fmaf(-su, s, x) > 0.0f) ? su : r;
    return r;
}

template <bool RP2, bool G2B>
__device__ __forceinline__ float term_fast(float Gx, float Gy, float Gz, const TermIn &a, float R, float invR,
                                           bool &bad) {
    float g2 = Gx * Gx + Gy * Gy;
    g2 = g2 + Gz * Gz;
    const float GMag = sqrt_core(g2);
    float num = Gx * a.rx + Gy * a.ry;
    num = num + Gz * a.rz;
    const bool dv = num > 0.0f;
    float tn = Gx * a.dx + Gy * a.dy;
    tn = tn + Gz * a.dz;
    const float tt = tn / (GMag * GMag);
    const float ex = a.x0 - (Gx * tt + a.Xc), ey = a.y0 - (Gy * tt + a.Yc), ez = a.z0 - (Gz * tt + a.Zc);
    float d2 = ex * ex + ey * ey;
    d2 = d2 + ez * ez;
    float Dk = sqrt_core(d2);
    Dk = 1.0f - (RP2 ? Dk * invR : Dk / R);
    Dk = Dk * Dk;
    // infinite or NaN d2 come out of sqrt_core as sqrtf gives them
    bad = !(g2 >= 0x1p-96f) || (!G2B && !(g2 <= 0x1p100f)) || !(d2 >= 0x1p-96f) || (dv && num < 0x1p-60f);
    return dv ? -Dk : Dk;
}

template <typename Tin> __device__ __forceinline__ float load_px(const Tin *p, long long i);
template <> __device__ __forceinline__ float load_px<uint16_t>(const uint16_t *p, long long i) {
    return (float)p[i];
}
template <> __device__ __forceinline__ float load_px<float>(const float *p, long long i) {
    return p[i];
}

__global__ void srrf_reset(int *counter, int *err) {
    if (threadIdx.x == 0) {
        *counter = 0;
        *err = 0;
    }
}

// Caller post-processing of the SR volume (multispot5_place_methods.cc:260-269):
// NaN -> 0, negative -> 0, in place, and the running maximum starting at 1.
// Non-negative floats order like their bit patterns, so the maximum is an
// unsigned atomic max of the bits (one per workgroup).
__global__ void srrf_clamp_init(unsigned *maxbits) {
    if (threadIdx.x == 0) *maxbits = __float_as_uint(1.0f);
}

__global__ void srrf_clamp(float *__restrict__ d, size_t n, unsigned *__restrict__ maxbits) {
    __shared__ float red[4];
    float mx = 1.0f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float v = d[i];
        if (v != v) v = 0.0f;
        if (v < 0.0f) v = 0.0f;
        d[i] = v;
        mx = v > mx ? v : mx;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float u = __shfl_xor(mx, o);
        mx = u > mx ? u : mx;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = red[w] > m ? red[w] : m;
        atomicMax(maxbits, __float_as_uint(m));
    }
}

// SR values of a region's SR pixels (multispot5_place_methods.cc:293-317):
// low-res pixel i expands to its M^3 SR pixels, x offset outermost and z
// offset innermost; value = SR volume at (x M + j1, y M + j2, z M + j3).
__global__ void srrf_gather_region(const float *__restrict__ sr, int X, int Y, int M, const int *__restrict__ px,
                                   long long n, float *__restrict__ out) {
    const int M3 = M * M * M;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const long long p = i / M3;
        const int j = (int)(i - p * M3);
        const int j1 = j / (M * M), j2 = (j / M) % M, j3 = j % M;
        const long long xs = (long long)px[3 * p] * M + j1, ys = (long long)px[3 * p + 1] * M + j2,
                        zs = (long long)px[3 * p + 2] * M + j3;
        out[i] = sr[zs * X * Y * M * M + ys * X * M + xs];
    }
}

// Plain streaming copy of n 16-byte vectors: the reference point for HBM
// byte counters (each byte read once and written once).
__global__ void copy_probe(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

template <typename Tin, int TZ, bool RP2>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3, 8))) void srrf_fused(LaunchParams P) {
    extern __shared__ __align__(16) float lds[];
    __shared__ int s_tile;
    const LaunchGeom &g = P.g;
    float *raw = lds + g.off_raw;
    float *G = lds + g.off_g;
    float *Px = lds + g.off_px;
    float *Py = lds + g.off_py;
    float *xtab = lds + g.off_xtab;  // kTX x {b, w0..w3}
    float *ztab = lds + g.off_ztab;  // TZ x noz x {wrel, wz0..3, valid, z0, dz}
    float *zcen = lds + g.off_zc;    // TZ: Zc
    float *ctab = lds + g.off_ctab;  // TZ x {wrel, wz0..3, valid} (voxel centre)
    float *dsum = lds + g.off_dsum;  // TZ x 256: per-voxel signed sums (thread-private)

    const int tid = threadIdx.x;
    // Each wave covers an 8 x 8 patch of the tile's columns, so that the
    // padding columns past a block's x edge fill whole waves (which skip).
    static_assert(kTX == 16 && kTY == 16 && kThreads == 256, "8 x 8 wave patches of a 16 x 16 tile");
    const int xl = ((tid >> 6) & 1) * 8 + (tid & 7), yl = (tid >> 7) * 8 + ((tid >> 3) & 7);
    const DevSamples *__restrict__ S = P.smp;
    // Sample tables are read in the innermost loops with uniform indices:
    // keep them in LDS (global reads there would wait on memory every
    // sample) and the scalars in registers.
    float *st_ox = lds + g.off_stab;
    float *st_oy = st_ox + g.st_oy;
    float *st_oz = st_ox + g.st_oz;
    int *st_grp = (int *)(st_ox + g.st_grp);  // ix, run0, nrun, ext
    int *st_run = (int *)(st_ox + g.st_run);  // start, len, iy, rx, ry, shift
    int *st_smp = (int *)(st_ox + g.st_smp);  // 32 iz (byte offset of the z-table entry), rz
    const float sR = S->R, sInvR = S->invR, sDiv = S->divisor;
    // merged sample groups when the plan chose them (two x offsets one
    // column apart served by one x pass), else one group per x offset run
    const bool merge = P.merge != 0;
    const int ngroups = merge ? S->nmgroups : S->ngroups;
    const int n_ox = S->nox, n_oy = S->noy;
    const bool fast_ok = S->fast_ok != 0;
    const float oxmin = S->oxmin, oxmax = S->oxmax, oymin = S->oymin, oymax = S->oymax;
    const float ozmin = S->ozmin, ozmax = S->ozmax;
    {
        const int nox = S->nox, noy = S->noy, nzz = S->noz, nrun = S->nruns, ns = S->n;
        for (int e = tid; e < nox; e += kThreads) st_ox[e] = S->ox[e];
        for (int e = tid; e < noy; e += kThreads) st_oy[e] = S->oy[e];
        for (int e = tid; e < nzz; e += kThreads) st_oz[e] = S->oz[e];
        for (int e = tid; e < ngroups; e += kThreads) {
            st_grp[4 * e] = merge ? S->mgrp_ix[e] : S->grp_ix[e];
            st_grp[4 * e + 1] = merge ? S->mgrp_run0[e] : S->grp_run0[e];
            st_grp[4 * e + 2] = merge ? S->mgrp_nrun[e] : S->grp_nrun[e];
            st_grp[4 * e + 3] = merge ? S->mgrp_ext[e] : 0;
        }
        for (int e = tid; e < nrun; e += kThreads) {
            const int s0 = S->run_start[e];
            st_run[6 * e] = s0;
            st_run[6 * e + 1] = S->run_len[e];
            st_run[6 * e + 2] = S->run_iy[e];
            st_run[6 * e + 3] = __float_as_int(S->rx[s0]);
            st_run[6 * e + 4] = __float_as_int(S->ry[s0]);
            st_run[6 * e + 5] = merge ? S->run_shift[e] : 0;
        }
        for (int e = tid; e < ns; e += kThreads) {
            st_smp[2 * e] = S->siz[e] * 8 * (int)sizeof(float);  // byte offset of its z-table entry
            st_smp[2 * e + 1] = __float_as_int(S->rz[e]);
        }
    }
    // end of sample-table setup
    float *scr = P.scratch + (long long)blockIdx.x * P.scratch_stride;
    const int NU = g.NU, NV = g.NV, NW = g.NW, TXP = g.TXP, noz = g.noz;
    const bool pxv = g.pxv != 0;
    const int NWP = g.NWP, VS = g.VS;
    const int RU = g.RU, RV = g.RV;

    for (;;) {
        __syncthreads();
        if (tid == 0) s_tile = atomicAdd(P.counter, 1);
        __syncthreads();
        const int ti = __builtin_amdgcn_readfirstlane(s_tile);
        if (ti >= P.ntiles) break;
        const DevTile tl = P.tiles[ti];
        const DevItem it = P.items[tl.item];
        const int M = it.M, W = it.W, H = it.H, D = it.D, T = it.T;
        const int X0 = tl.X0, Y0 = tl.Y0, Z0 = tl.Z0;

        // Low-res extents of all taps of this tile (monotone in coordinate).
        const int ulo = first_tap(((float)X0 + 0.5f) + oxmin, M);
        const int vlo = first_tap(((float)Y0 + 0.5f) + oymin, M);
        const int wlo = first_tap(((float)Z0 + 0.5f) + ozmin, M);
        const int uhi = first_tap(((float)(X0 + kTX - 1) + 0.5f) + oxmax, M) + 3;
        const int vhi = first_tap(((float)(Y0 + kTY - 1) + 0.5f) + oymax, M) + 3;
        const int whi = first_tap(((float)(Z0 + TZ - 1) + 0.5f) + ozmax, M) + 3;
        if (uhi - ulo + 1 > NU || vhi - vlo + 1 > NV || whi - wlo + 1 > NW) {
            if (tid == 0) atomicOr(P.err, 1);
            continue;
        }
        // This tile's own extents (a batch plans LDS for its largest tile;
        // tiles of larger magnification cover fewer low-res rows).  Every tap
        // of the tile lies inside them; strides stay the planned ones.
        const int nu = uhi - ulo + 1, nv = vhi - vlo + 1, nw = whi - wlo + 1;
        const int ru = nu + 2 * g.hx, rv_ = nv + 2 * g.hy, rw = nw + 2 * g.hz;
        const int nr_raw = ru * rv_ * rw;
        const float r_ru = 1.0f / (float)ru, r_rv = 1.0f / (float)rv_, r_nu = 1.0f / (float)nu;
        const float r_nv = 1.0f / (float)nv, r_nw = 1.0f / (float)nw;

        // z tables (uniform per tile, frame independent).
        for (int e = tid; e < TZ * noz; e += kThreads) {
            const int Z = e / noz, iz = e % noz;
            const float Zc = (float)(Z0 + Z) + 0.5f;
            const float z0 = Zc + st_oz[iz];
            int b, valid;
            float w0, w1, w2, w3;
            axis_weights(z0, M, D, b, w0, w1, w2, w3, valid);
            // An out-of-range sample (interpolateG returns 0, :146-147) gets
            // zero weights on the axis that fails: its interpolated gradient
            // is then +-0, which takes the same path as an exact 0 (g2 = 0:
            // the general term, NaN).  Likewise for x and y below.
            if (!valid) w0 = w1 = w2 = w3 = 0.0f;
            float *zt = ztab + e * 8;
            // first tap row as the float offset of its y-interpolated
            // columns, Py[(w 3 + c) 256 + tid]
            zt[0] = __int_as_float(clampi(b - wlo, 0, NW - 4) * 3 * kThreads);
            zt[1] = w0; zt[2] = w1; zt[3] = w2; zt[4] = w3;
            zt[5] = __int_as_float(valid);
            zt[6] = z0;
            zt[7] = z0 - Zc;
        }
        if (tid < TZ) {
            const float Zc = (float)(Z0 + tid) + 0.5f;
            zcen[tid] = Zc;
            int b, valid;
            float w0, w1, w2, w3;
            axis_weights(Zc, M, D, b, w0, w1, w2, w3, valid);
            float *ct = ctab + tid * 6;
            ct[0] = __int_as_float(clampi(b - wlo, 0, NW - 4));
            ct[1] = w0; ct[2] = w1; ct[3] = w2; ct[4] = w3;
            ct[5] = __int_as_float(valid);
        }

        const int Xs = X0 + xl, Ys = Y0 + yl;
        const float Xc = (float)Xs + 0.5f, Yc = (float)Ys + 0.5f;

        // Voxels whose value needs no sample loop: a sample out of range on
        // any axis makes every frame's radiality NaN (its gradient is zero:
        // the general term, NaN, and a sum with NaN), and voxels past the
        // block are not stored.  zmask: tile slices to skip; colbad: this
        // thread's column.  The range test is axis_weights' own.
        __syncthreads();  // z tables
        unsigned zmask = 0;
        for (int Z = 0; Z < TZ; Z++) {
            bool bad = (Z0 + Z) < it.bz0 || (Z0 + Z) >= it.bz1;
            for (int iz = 0; iz < noz; iz++) bad = bad || !__float_as_int(ztab[(Z * noz + iz) * 8 + 5]);
            if (bad) zmask |= 1u << Z;
        }
        bool colbad = Xs < it.bx0 || Xs >= it.bx1 || Ys < it.by0 || Ys >= it.by1;
        for (int i = 0; i < n_ox; i++) {
            const float xx = (Xc + st_ox[i]) / (float)M;
            colbad = colbad || xx < 1.5f || xx > (float)W - 1.5f;
        }
        for (int i = 0; i < n_oy; i++) {
            const float yy = (Yc + st_oy[i]) / (float)M;
            colbad = colbad || yy < 1.5f || yy > (float)H - 1.5f;
        }
        const unsigned zall = (TZ >= 32) ? ~0u : ((1u << TZ) - 1u);
        const bool tile_skip = __syncthreads_and(colbad || zmask == zall) != 0;
        // An idle tile: no voxel has its samples in range, so every frame's
        // radiality is NaN (0 after RemoveCon).  With integer input (finite,
        // non-negative centre intensities) or without intensity weighting
        // (intensities stay 0) the temporal modes then give the same outputs
        // as with all intensities 0: NaN * I = NaN, 0 * I = +0, and TRPPM's
        // pair sum of zero series is +0 before its intensity factor.  Such a
        // tile skips the frame body (halo, gradient, intensities, samples and
        // their barriers) and only accumulates that constant frame value.
        const bool idle = tile_skip && (kIntInput<Tin> || P.iw != 1);

        float acc0[TZ], acc1[TZ];
#pragma unroll
        for (int z = 0; z < TZ; z++) { acc0[z] = 0.0f; acc1[z] = 0.0f; }

        // raw halo (clamp-to-edge inside the stack, kernel.cu:59-80)
        const Tin *src = (const Tin *)it.src;
        const int rx0 = ulo - g.hx, ry0 = vlo - g.hy, rz0 = wlo - g.hz;

        for (int t = 0; t < T; t++) {
            float Iv[TZ], rv[TZ];
            if (idle) {
                // the frame value of every voxel (NaN, as for the marked voxels
                // below) and intensities 0: no halo, gradient or samples, and no
                // barrier (idle is the same for the whole workgroup)
#pragma unroll
                for (int Z = 0; Z < TZ; Z++) {
                    float dd = __builtin_nanf("");
                    if (P.renorm == 1) dd = 0.5f + (dd / 2.0f);
                    rv[Z] = (P.remove_con == 1) ? fmaxf(dd, 0.0f) : dd;
                    Iv[Z] = 0.0f;
                }
            } else {
                __syncthreads();
                {
                    const Tin *srct = src + (long long)t * it.st;
                    for (int e = tid; e < nr_raw; e += kThreads) {
                        const int q = qdiv(e, r_ru), c = qdiv(q, r_rv);
                        const int a = e - q * ru, bb = q - c * rv_;
                        const int x = clampi(rx0 + a, 0, W - 1) + it.ox;
                        const int y = clampi(ry0 + bb, 0, H - 1) + it.oy;
                        const int z = clampi(rz0 + c, 0, D - 1) + it.oz;
                        raw[(c * RV + bb) * RU + a] = load_px<Tin>(srct, (long long)x + y * it.sy + z * it.sz);
                    }
                }
                __syncthreads();

                // ---- gradient planes G[c][w][v][u] (only the sample passes read them)
                // Integer input with sums of at most 256 terms (P.sep, set by the
                // host): every partial sum is an integer below 2^24, exact in FP32
                // in any order, so the smoothed gradient is formed from shared box
                // sums -- Bz (z sums) for the x and y derivatives, By (y sums) for
                // the z derivative -- and equals the term-by-term sum bit for bit.
                // Bz and By live in the x-plane area, which the intensity and x
                // passes overwrite only after the next barrier.
                const bool sep = kIntInput<Tin> && P.sep && P.smooth != 0 && !tile_skip;
                if (sep) {
                    float *Bz = Px;                  // [nw][rv_][ru]
                    float *By = Px + ru * rv_ * nw;  // [rw][nv][ru]
                    const int nbz = ru * rv_ * nw, nby = ru * nv * rw;
                    const int mR = P.mR, xR = P.xR, yR = P.yR, zR = P.zR;
                    for (int e = tid; e < nbz + nby; e += kThreads) {
                        float s = 0.0f;
                        if (e < nbz) {
                            const int q = qdiv(e, r_ru), w = qdiv(q, r_rv);
                            const int a = e - q * ru, b = q - w * rv_;
                            const float *r = raw + ((w + g.hz) * RV + b) * RU + a;
                            for (int k = -zR; k <= zR; k++) s += r[k * RV * RU];
                            Bz[e] = s;
                        } else {
                            const int f = e - nbz;
                            const int q = qdiv(f, r_ru), c = qdiv(q, r_nv);
                            const int a = f - q * ru, v = q - c * nv;
                            const float *r = raw + (c * RV + v + g.hy) * RU + a;
                            for (int j = -yR; j <= yR; j++) s += r[j * RU];
                            By[f] = s;
                        }
                    }
                    __syncthreads();
                    const int NP = NU * NV * NW;
                    for (int e = tid; e < nu * nv * nw; e += kThreads) {
                        const int q = qdiv(e, r_nu), w = qdiv(q, r_nv);
                        const int u = e - q * nu, v = q - w * nv;
                        const int cu = u + g.hx, cv = v + g.hy, cw = w + g.hz;
                        const float *bz = Bz + w * rv_ * ru;
                        float gx = 0.0f, gy = 0.0f, gz = 0.0f;
                        for (int i = 1; i <= mR; i++) {
                            float sp = 0.0f, sm = 0.0f;
                            for (int j = -yR; j <= yR; j++) {
                                sp += bz[(cv + j) * ru + cu + i];
                                sm += bz[(cv + j) * ru + cu - i];
                            }
                            gx += sp - sm;
                        }
                        for (int j = 1; j <= mR; j++) {
                            float sp = 0.0f, sm = 0.0f;
                            for (int i = -xR; i <= xR; i++) {
                                sp += bz[(cv + j) * ru + cu + i];
                                sm += bz[(cv - j) * ru + cu + i];
                            }
                            gy += sp - sm;
                        }
                        for (int k = 1; k <= mR; k++) {
                            float sp = 0.0f, sm = 0.0f;
                            for (int i = -xR; i <= xR; i++) {
                                sp += By[((cw + k) * nv + v) * ru + cu + i];
                                sm += By[((cw - k) * nv + v) * ru + cu + i];
                            }
                            gz += sp - sm;
                        }
                        const int gi3 = (w * NV + v) * NU + u;
                        G[gi3] = gx;
                        G[NP + gi3] = gy;
                        G[2 * NP + gi3] = gz;
                    }
                }
                for (int e = tid; e < (tile_skip || sep ? 0 : nu * nv * nw); e += kThreads) {
                    const int q = qdiv(e, r_nu), w = qdiv(q, r_nv);
                    const int u = e - q * nu, v = q - w * nv;
                    const int gi3 = (w * NV + v) * NU + u;
                    const int cu = u + g.hx, cv = v + g.hy, cw = w + g.hz;
                    auto R3 = [&](int du, int dv, int dw) {
                        return raw[((cw + dw) * RV + (cv + dv)) * RU + (cu + du)];
                    };
                    float gx, gy, gz;
                    if (P.smooth == 0) {
                        gx = -R3(-1, 0, 0) + R3(1, 0, 0);
                        gy = -R3(0, -1, 0) + R3(0, 1, 0);
                        gz = -R3(0, 0, -1) + R3(0, 0, 1);
                    } else {
                        const int mR = P.mR, xR = P.xR, yR = P.yR, zR = P.zR;
                        float s = 0.0f;
                        for (int i = -mR; i <= mR; i++)
                            for (int j = -yR; j <= yR; j++)
                                for (int k = -zR; k <= zR; k++) {
                                    if (i < 0) s -= R3(i, j, k);
                                    else if (i > 0) s += R3(i, j, k);
                                }
                        gx = s;
                        s = 0.0f;
                        for (int j = -mR; j <= mR; j++)
                            for (int i = -xR; i <= xR; i++)
                                for (int k = -zR; k <= zR; k++) {
                                    if (j < 0) s -= R3(i, j, k);
                                    else if (j > 0) s += R3(i, j, k);
                                }
                        gy = s;
                        s = 0.0f;
                        for (int k = -mR; k <= mR; k++)
                            for (int i = -xR; i <= xR; i++)
                                for (int j = -yR; j <= yR; j++) {
                                    if (k < 0) s -= R3(i, j, k);
                                    else if (k > 0) s += R3(i, j, k);
                                }
                        gz = s;
                    }
                    const int NP = NU * NV * NW;
                    G[gi3] = gx;
                    G[NP + gi3] = gy;
                    G[2 * NP + gi3] = gz;
                }

                // ---- intensity at the voxel centre (interpolateIntensity :257-287),
                // before the sample passes: the raw halo is dead afterwards and
                // its LDS is reused by the y-interpolated columns.
#pragma unroll
                for (int Z = 0; Z < TZ; Z++) Iv[Z] = 0.0f;
                if (P.iw == 1) {
                    if (tid < kTX) {
                        int b, valid;
                        float w0, w1, w2, w3;
                        axis_weights((float)(X0 + tid) + 0.5f, M, W, b, w0, w1, w2, w3, valid);
                        float *xt = xtab + tid * 6;
                        xt[0] = __int_as_float(clampi(b - ulo, 0, NU - 4));
                        xt[1] = w0; xt[2] = w1; xt[3] = w2; xt[4] = w3;
                    }
                    __syncthreads();
                    for (int e = tid; e < nv * nw * kTX; e += kThreads) {
                        const int q = e / kTX, v = qdiv(q, r_nw);
                        const int X = e % kTX, w = q - v * nw;
                        const float *xt = xtab + X * 6;
                        const int ub = __float_as_int(xt[0]);
                        const float *rr = raw + ((w + g.hz) * RV + (v + g.hy)) * RU + ub + g.hx;
                        float p = 0.0f;
                        p = p + rr[0] * xt[1];
                        p = p + rr[1] * xt[2];
                        p = p + rr[2] * xt[3];
                        p = p + rr[3] * xt[4];
                        Px[pxv ? v * VS + X * NWP + w : (v * NW + w) * TXP + X] = p;
                    }
                    __syncthreads();
                    int vb, validy;
                    float wy0, wy1, wy2, wy3;
                    axis_weights(Yc, M, H, vb, wy0, wy1, wy2, wy3, validy);
                    const int vrel = clampi(vb - vlo, 0, NV - 4);
                    const int validx = !((Xc / (float)M) < 1.5f || (Xc / (float)M) > (float)W - 1.5f);
                    for (int w = 0; w < nw; w++) {
                        const float *pp = Px + (pxv ? vrel * VS + xl * NWP + w : (vrel * NW + w) * TXP + xl);
                        const int vs = pxv ? VS : NW * TXP;
                        float q = 0.0f;
                        q = q + pp[0] * wy0;
                        q = q + pp[vs] * wy1;
                        q = q + pp[2 * vs] * wy2;
                        q = q + pp[3 * vs] * wy3;
                        Py[w * kThreads + tid] = q;
                    }
#pragma unroll
                    for (int Z = 0; Z < TZ; Z++) {
                        const float *ct = ctab + Z * 6;
                        const int wrel = __float_as_int(ct[0]);
                        const float *qq = Py + wrel * kThreads + tid;
                        float sv = 0.0f;
                        sv = sv + qq[0] * ct[1];
                        sv = sv + qq[kThreads] * ct[2];
                        sv = sv + qq[2 * kThreads] * ct[3];
                        sv = sv + qq[3 * kThreads] * ct[4];
                        const int valid = validx & validy & __float_as_int(ct[5]);
                        Iv[Z] = valid ? fmaxf(sv, 0.0f) : 0.0f;
                    }
                }

                for (int z = 0; z < TZ; z++) dsum[z * kThreads + tid] = 0.0f;

                for (int gi = 0; gi < ngroups && !tile_skip; gi++) {
                    const int ix = st_grp[4 * gi];
                    const float oxv = st_ox[ix];
                    // a merged group's planes have one more column (read by its
                    // shifted runs)
                    const int xw = kTX + st_grp[4 * gi + 3];
                    const float r_xw = 1.0f / (float)xw;
                    if (tid < xw) {
                        int b, valid;
                        float w0, w1, w2, w3;
                        axis_weights(((float)(X0 + tid) + 0.5f) + oxv, M, W, b, w0, w1, w2, w3, valid);
                        if (!valid) w0 = w1 = w2 = w3 = 0.0f;
                        float *xt = xtab + tid * 6;
                        xt[0] = __int_as_float(clampi(b - ulo, 0, NU - 4));
                        xt[1] = w0; xt[2] = w1; xt[3] = w2; xt[4] = w3;
                    }
                    __syncthreads();
                    // ---- x pass: Px[c][v][w][X] = sum_i G[c][w][v][b+i] * wx_i ----
                    {
                        const int NP = NU * NV * NW;
                        // element e = (v, w, X): the constant divisor kTX for a
                        // plain group, the reciprocal quotient for kTX + 1 columns
                        auto xelem = [&](int X, int q) {
                            const int v = qdiv(q, r_nw), w = q - v * nw;
                            const float *xt = xtab + X * 6;
                            const int ub = __float_as_int(xt[0]);
                            const float w0 = xt[1], w1 = xt[2], w2 = xt[3], w3 = xt[4];
                            const int gb = (w * NV + v) * NU + ub;
#pragma unroll
                            for (int c = 0; c < 3; c++) {
                                const float *gg = G + c * NP + gb;
                                float p = gg[0] * w0;
                                p = p + gg[1] * w1;
                                p = p + gg[2] * w2;
                                p = p + gg[3] * w3;
                                Px[pxv ? (c * NV + v) * VS + X * NWP + w : ((c * NV + v) * NW + w) * TXP + X] = p;
                            }
                        };
                        if (xw == kTX) {
                            for (int e = tid; e < nv * nw * kTX; e += kThreads) xelem(e % kTX, e / kTX);
                        } else {
                            for (int e = tid; e < nv * nw * xw; e += kThreads) {
                                const int q = qdiv(e, r_xw);
                                xelem(e - q * xw, q);
                            }
                        }
                    }
                    __syncthreads();

                    const int r0 = st_grp[4 * gi + 1], nr = st_grp[4 * gi + 2];
                    // a wave whose voxels all skip does no y pass or samples
                    const bool wave_skip = __all(colbad || zmask == zall);
                    for (int r = r0; r < r0 + nr && !wave_skip; r++) {
                        const int iy = st_run[6 * r + 2];
                        // the run's own x offset (a shifted run: the next plane column)
                        const float rxs = __int_as_float(st_run[6 * r + 3]);
                        const int xs = xl + st_run[6 * r + 5];
                        const float x0 = Xc + rxs;
                        const float dxs = x0 - Xc;
                        const float y0 = Yc + st_oy[iy];
                        const float dys = y0 - Yc;
                        int vb, validy;
                        float wy0, wy1, wy2, wy3;
                        axis_weights(y0, M, H, vb, wy0, wy1, wy2, wy3, validy);
                        if (!validy) wy0 = wy1 = wy2 = wy3 = 0.0f;
                        const int vrel = clampi(vb - vlo, 0, NV - 4);
                        // ---- y pass: Py[w][c][tid] = sum_j Px[c][vrel+j][w][X] * wy_j
                        if (pxv) {
                            // column layout: the 3 x 4 tap rows of two w values
                            // are loaded as pairs first, then summed in order
                            const float *pb = Px + vrel * VS + xs * NWP;
#pragma unroll 4
                            for (int w = 0; w < nw; w += 2) {
                                float2 pv[3][4];
#pragma unroll
                                for (int c = 0; c < 3; c++)
#pragma unroll
                                    for (int j = 0; j < 4; j++)
                                        pv[c][j] = *(const float2 *)(pb + (c * NV + j) * VS + w);
#pragma unroll
                                for (int c = 0; c < 3; c++) {
                                    float q = pv[c][0].x * wy0;
                                    q = q + pv[c][1].x * wy1;
                                    q = q + pv[c][2].x * wy2;
                                    q = q + pv[c][3].x * wy3;
                                    Py[(w * 3 + c) * kThreads + tid] = q;
                                }
                                if (w + 1 < nw) {
#pragma unroll
                                    for (int c = 0; c < 3; c++) {
                                        float q = pv[c][0].y * wy0;
                                        q = q + pv[c][1].y * wy1;
                                        q = q + pv[c][2].y * wy2;
                                        q = q + pv[c][3].y * wy3;
                                        Py[((w + 1) * 3 + c) * kThreads + tid] = q;
                                    }
                                }
                            }
                        } else {
                            for (int w = 0; w < nw; w++) {
#pragma unroll
                                for (int c = 0; c < 3; c++) {
                                    const float *pp = Px + ((c * NV + vrel) * NW + w) * TXP + xs;
                                    const int vs = NW * TXP;
                                    float q = pp[0] * wy0;
                                    q = q + pp[vs] * wy1;
                                    q = q + pp[2 * vs] * wy2;
                                    q = q + pp[3 * vs] * wy3;
                                    Py[(w * 3 + c) * kThreads + tid] = q;
                                }
                            }
                        }
                        const int s0 = st_run[6 * r], ns = st_run[6 * r + 1];
                        const float rys = __int_as_float(st_run[6 * r + 4]);
                        // Voxel pairs (Z, Z+1): two independent chains per sample.
                        for (int Z = 0; Z < TZ; Z += 2) {
                            if (__all(colbad || ((zmask >> Z) & 3u) == 3u)) continue;
                            float dd0 = dsum[Z * kThreads + tid], dd1 = dsum[(Z + 1) * kThreads + tid];
                            TermIn a0, a1;
                            a0.rx = a1.rx = rxs; a0.ry = a1.ry = rys;
                            a0.dx = a1.dx = dxs; a0.dy = a1.dy = dys;
                            a0.x0 = a1.x0 = x0; a0.y0 = a1.y0 = y0;
                            a0.Xc = a1.Xc = Xc; a0.Yc = a1.Yc = Yc;
                            a0.Zc = zcen[Z];
                            a1.Zc = zcen[Z + 1];
                            const float *zrow0 = ztab + Z * noz * 8, *zrow1 = zrow0 + noz * 8;
                            const float *Pyt = Py + tid;
                            for (int s = s0; s < s0 + ns; s++) {
                                const int izb = st_smp[2 * s];
                                a0.rz = a1.rz = __int_as_float(st_smp[2 * s + 1]);
                                const float *zt0 = (const float *)((const char *)zrow0 + izb);
                                const float *zt1 = (const float *)((const char *)zrow1 + izb);
                                const int wr0 = __float_as_int(zt0[0]), wr1 = __float_as_int(zt1[0]);
                                a0.z0 = zt0[6]; a0.dz = zt0[7];
                                a1.z0 = zt1[6]; a1.dz = zt1[7];
                                // ---- z pass (interpolateG :154-167, innermost sum last)
                                float G0[3], G1[3];
#pragma unroll
                                for (int c = 0; c < 3; c++) {
                                    const float *q0 = Pyt + wr0 + c * kThreads;
                                    const float *q1 = Pyt + wr1 + c * kThreads;
                                    float s0v = q0[0] * zt0[1];
                                    float s1v = q1[0] * zt1[1];
                                    s0v = s0v + q0[3 * kThreads] * zt0[2];
                                    s1v = s1v + q1[3 * kThreads] * zt1[2];
                                    s0v = s0v + q0[6 * kThreads] * zt0[3];
                                    s1v = s1v + q1[6 * kThreads] * zt1[3];
                                    s0v = s0v + q0[9 * kThreads] * zt0[4];
                                    s1v = s1v + q1[9 * kThreads] * zt1[4];
                                    G0[c] = s0v;
                                    G1[c] = s1v;
                                }
                                bool b0, b1;
                                float t0 = term_fast<RP2, kIntInput<Tin>>(G0[0], G0[1], G0[2], a0, sR, sInvR, b0);
                                float t1 = term_fast<RP2, kIntInput<Tin>>(G1[0], G1[1], G1[2], a1, sR, sInvR, b1);
                                b0 = b0 || !fast_ok;
                                b1 = b1 || !fast_ok;
                                if (b0 || b1) {
                                    if (b0) t0 = term_general<RP2>(G0[0], G0[1], G0[2], a0, sR, sInvR);
                                    if (b1) t1 = term_general<RP2>(G1[0], G1[1], G1[2], a1, sR, sInvR);
                                }
                                dd0 = dd0 + t0;
                                dd1 = dd1 + t1;
                            }
                            dsum[Z * kThreads + tid] = dd0;
                            dsum[(Z + 1) * kThreads + tid] = dd1;
                        }
                    }
                    __syncthreads();
                }


                // ---- per-frame radiality value (kernel.cu:244-251)
#pragma unroll
                for (int Z = 0; Z < TZ; Z++) {
                    float dd = (colbad || ((zmask >> Z) & 1u)) ? __builtin_nanf("") : dsum[Z * kThreads + tid] / sDiv;
                    if (P.renorm == 1) dd = 0.5f + (dd / 2.0f);
                    rv[Z] = (P.remove_con == 1) ? fmaxf(dd, 0.0f) : dd;
                }
            }

            // ---- temporal accumulation (same order as the reference)
            const float Tf = (float)T;
#pragma unroll
            for (int Z = 0; Z < TZ; Z++) {
                const float r = rv[Z];
                if (P.method == 0) {
                    acc0[Z] = fmaxf(P.iw == 1 ? r * Iv[Z] : r, acc0[Z]);
                } else if (P.method == 1) {
                    acc0[Z] += (P.iw == 1) ? r * Iv[Z] : r;
                } else if (P.method == 2) {
                    if (P.iw == 1) acc0[Z] += Iv[Z] / Tf;
                    scr[((long long)t * TZ + Z) * kThreads + tid] = r;
                } else {
                    acc0[Z] += r * Iv[Z];
                    acc1[Z] += r;
                    scr[((long long)t * TZ + Z) * kThreads + tid] = r;
                }
            }
        }  // frames

        // ---- final temporal reduction and output
        const bool colok = Xs >= it.bx0 && Xs < it.bx1 && Ys >= it.by0 && Ys < it.by1;
        const float Tf = (float)T;
        // park the register accumulators in the (now free) private LDS rows
        float *acc_l = dsum;
        float *acc1_l = Py;  // 3*NW*256 >= 2*TZ*256 is checked by the planner
#pragma unroll
        for (int Z = 0; Z < TZ; Z++) {
            acc_l[Z * kThreads + tid] = acc0[Z];
            acc1_l[Z * kThreads + tid] = acc1[Z];
        }
        for (int Z = 0; Z < TZ; Z++) {
            const int Zs = Z0 + Z;
            const float a0 = acc_l[Z * kThreads + tid], a1 = acc1_l[Z * kThreads + tid];
            float out = 0.0f;
            auto ser = [&](int tt) -> float {
                return scr[((long long)tt * TZ + Z) * kThreads + tid];
            };
            if (P.method == 0) {
                out = a0;
            } else if (P.method == 1) {
                out = a0 / Tf;
            } else if (P.method == 2) {
                // calculatePairwiseProductSum :394-416
                float pps = 0.0f;
                int counter = 0;
                for (int t0 = 0; t0 < T; t0++) {
                    const float r0 = fmaxf(ser(t0), 0.0f);
                    if (r0 > 0) {
                        for (int t1 = t0; t1 < T; t1++) {
                            const float r1 = fmaxf(ser(t1), 0.0f);
                            pps += r0 * r1;
                            counter++;
                        }
                    } else
                        counter += T - t0;
                }
                pps /= fmaxf((float)counter, 1.0f);
                if (P.iw == 1) pps *= a0;
                out = pps;
            } else {
                // calculateACRF :458-554
                const int ord = P.order;
                const float mean = a1 / Tf;
                const float IW = a0 / Tf;
                auto rr = [&](int tt) -> float { return tt < T ? ser(tt) : 0.0f; };
                float ABCD = 0, ABC = 0, AB = 0, CD = 0, AC = 0, BD = 0, AD = 0, BC = 0;
                float acc = 0.0f;
                if (P.itc != 1) {
                    for (int tt = 0; tt < T - ord; tt++) {
                        const float A = rr(tt) - mean, B = rr(tt + 1) - mean;
                        AB += (A * B);
                        if (ord == 3) ABC += (A * B * (rr(tt + 2) - mean));
                        if (ord == 4) {
                            const float C = rr(tt + 2) - mean, Dd = rr(tt + 3) - mean;
                            ABCD += A * B * C * Dd;
                            CD += C * Dd; AC += A * C; BD += B * Dd; AD += A * Dd; BC += B * C;
                        }
                    }
                    if (ord == 3) acc = fabsf(ABC) / Tf;
                    else if (ord == 4) acc = fabsf(ABCD - AB * CD - AC * BD - AD * BC) / Tf;
                    else acc = fabsf(AB) / Tf;
                } else {
                    int nB = T;
                    while (nB > ord) {
                        AB = 0.0f;
                        for (int tt = 0; tt < nB - ord; tt++) {
                            const int tBin = tt * ord;
                            const float A = rr(tt) - mean, B = rr(tt + 1) - mean;
                            AB += (A * B);
                            if (ord == 3) ABC += (A * B * (rr(tt + 2) - mean));
                            if (ord == 4) {
                                const float C = rr(tt + 2) - mean, Dd = rr(tt + 3) - mean;
                                ABCD += A * B * C * Dd;
                                CD += C * Dd; AC += A * C; BD += B * Dd; AD += A * Dd; BC += B * C;
                            }
                            float nv = 0.0f;
                            scr[((long long)tt * TZ + Z) * kThreads + tid] = 0.0f;
                            if (tBin < nB)
                                for (int q = 0; q < ord; q++) {
                                    nv += rr(tBin + q) / (float)ord;
                                    scr[((long long)tt * TZ + Z) * kThreads + tid] = nv;
                                }
                        }
                        if (ord == 3) acc += fabsf(ABC) / (float)nB;
                        else if (ord == 4) acc += fabsf((ABCD - AB * CD - AC * BD - AD * BC)) / (float)nB;
                        else acc += fabsf(AB) / (float)nB;
                        nB = nB / ord;
                    }
                }
                if (P.iw == 1) acc = IW * acc;
                out = acc;
            }
            if (colok && Zs >= it.bz0 && Zs < it.bz1)
                it.out[(long long)Zs * it.osz + (long long)Ys * it.osy + Xs] = out;
        }
    }
}

// Explicit instantiations used by the host launcher.
#define RFBA_INST(T, Z)                                                  \
    template __global__ void srrf_fused<T, Z, false>(LaunchParams);     \
    template __global__ void srrf_fused<T, Z, true>(LaunchParams);
RFBA_INST(uint16_t, 4)
RFBA_INST(uint16_t, 8)
RFBA_INST(uint16_t, 16)
RFBA_INST(float, 4)
RFBA_INST(float, 8)
RFBA_INST(float, 16)

}  // namespace rfba
