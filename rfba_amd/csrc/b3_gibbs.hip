// b3_gibbs.hip -- Gibbs sampling of the blink states of RFBA's 3B model on
// gfx950: GibbsSampler2::next (RFBA Main/sampled_multispot.h:681-709) for
// many chains at once, plus the two spot tables the sampler consumes,
// get_spot_pixels (multispot5.cc:1001-1020) and compute_spot_intensity
// (sampled_multispot.h:183-191).
//
// The masked on/off log-probabilities of spot k in frame t
// (spot_with_background.hh:120-147 over the mask M_k) differ by
// B_on - B_off = -d_t(k) / (2 var), with r0 the data minus the base minus
// the other spots that are on, each over its own mask (add_spot,
// sampled_multispot.h:124-131):
//   d_t(k) = sum_{M_k} (r0 - mu_k)^2 - r0^2 = S2_k - 2 (C_t(k) - A_t(k)),
//   S2_k = sum_{M_k} mu_k^2,   C_t(k) = sum_{M_k} mu_k (data_t - base),
//   A_t(k) = sum_{j != k on in frame t} O_kj,   O_kj = sum_{M_k and M_j} mu_k mu_j.
// Everything else in B is common to the three states of a frame and cancels
// in both the forward recursion and the backward sampling.  S2, C and O are
// computed once per plan on the device (FP64 sums over the masks); A is kept
// in FP64 and changed by +-O_kj in the frames where spot j turns on or off.
// A sweep therefore reads no pixels: per spot update, the frames' d_t, the
// forward recursion and backward sampling on one lane, and the A updates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "b3_internal.h"
#include "b3_wave.h"

namespace rfba3b {

constexpr int kSweepThreads = 64;   // one wave per chain
constexpr int kStatThreads = 256;   // statistics kernels
constexpr int kCpartThreads = 512;  // b3_gibbs_cpart: more waves per staged range (measured best)
constexpr int kOpartThreads = 256;  // b3_gibbs_opart (measured best)
constexpr int kRange = 1024;        // pixels per range of the statistics kernels
constexpr int kMtN = 624;           // MT19937 state words (mersenne.cpp MERS_N)

struct GibbsDev {
    int P, T, K;
    const float *data;    // [T][P]
    const double *spot;   // [K][P]
    const int *moff;      // [K + 1]
    const int *midx;      // mask pixel indices
    uint8_t *sample;      // [K][T]
    uint8_t *samples;     // [S][K][T] or null
    double *C;            // [K][T]
    double *A;            // [K][T]
    double *O;            // [K][K]
    double *S2;           // [K]
    double base;
    uint32_t *mt;         // [kMtN + 1]: state words, then the position of the next draw
};

// A chain as the statistics kernels see it (plan creation only).
struct StatDev {
    int P, T, K, ne;      // ne: list entries
    const float *data;    // [T][P]
    const double *spot;   // [K][P]
    const int *moff;      // [K + 1]
    const int *midx;      // mask pixel indices
    double base;
    const int *roff;      // [R + 1]: the list entries of range r
    const int *lst;       // [ne][4]: (spot, range, first mask entry, end)
    const int *koff;      // [K + 1]: the list entries of spot k in kent
    const int *kent;      // [ne]: list entries of each spot, in range order
    double *mu;           // [mask entries]: the spot's value at each mask pixel
    double *Cp;           // [ne][T]
    double *Op;           // [ne][K]
    double *C, *O, *S2;
};

struct StatTask {
    int chain, r, g;      // range, frame group
};

struct GibbsParams {
    const GibbsDev *chains;
    const int *order;     // launch order (largest work first)
    int samples, iters, Tcap, pad;
    const double *var;    // [samples]
    double A[9], pi[3];
};

struct SpotTask {
    int chain, k;
};

// ---------------------------------------------------------------------------
// MT19937 (mersenne.cpp BRandom / Random): tempering and the state update.

__device__ __forceinline__ uint32_t mt_temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

__device__ __forceinline__ uint32_t mt_word(uint32_t cur, uint32_t nxt, uint32_t far) {
    const uint32_t y = (cur & 0x80000000u) | (nxt & 0x7fffffffu);
    return far ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

// The sequential update new[i] = f(old[i], old[i+1], (i < 227 ? old : new)[i+397 mod 624])
// by one wave in four dependent ranges: [0,227) reads only old words,
// [227,454) and [454,623) read the new words 227 below them, and 623 reads
// new[0] and new[396].
__device__ void mt_twist64(uint32_t *mt, int lane) {
    uint32_t v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = lane + 64 * r;
        if (i < 227) v[r] = mt_word(mt[i], mt[i + 1], mt[i + 397]);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = lane + 64 * r;
        if (i < 227) mt[i] = v[r];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = 227 + lane + 64 * r;
        if (i < 454) v[r] = mt_word(mt[i], mt[i + 1], mt[i - 227]);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = 227 + lane + 64 * r;
        if (i < 454) mt[i] = v[r];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const int i = 454 + lane + 64 * r;
        if (i < 623) v[r] = mt_word(mt[i], mt[i + 1], mt[i - 227]);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const int i = 454 + lane + 64 * r;
        if (i < 623) mt[i] = v[r];
    }
    __syncthreads();
    if (lane == 0) mt[623] = mt_word(mt[623], mt[0], mt[396]);
    __syncthreads();
}

// select_random_element over weights w (forward_algorithm.h:502-513): the
// first state whose running sum reaches u * (w0 + w1 + w2), else the last.
__device__ __forceinline__ int select3(double w0, double w1, double w2, double u) {
    double sum = 0.0;
    sum += w0;
    sum += w1;
    sum += w2;
    const double choice = u * sum;
    double total = w0;
    if (choice <= total) return 0;
    total += w1;
    if (choice <= total) return 1;
    return 2;
}

// The transfer matrix of frame t (transfer_matrix) with x = B_on - B_off =
// -(S2 - 2 (C_t - A_t)) / (2 var); the identity past T.
__device__ __forceinline__ void frame_matrix(int t, int T, double S2, const double *Ck, const double *Ak, double i2v,
                                             const double *A, const double *pi, double (&Mm)[9], double (&Me)[9]) {
    if (t < T)
        transfer_matrix(-(S2 - 2.0 * (Ck[t] - Ak[t])) * i2v, t == 0, A, pi, Mm, Me);
    else
        xident(Mm, Me);
}

// select_random_element (forward_algorithm.h:502-513) over the extended
// weights w_i = m_i 2^e_i, aligned to the largest as sample_unscaled_log
// scales its weights by the largest (:520-532).
__device__ __forceinline__ int xselect3(double m0, double e0, double m1, double e1, double m2, double e2, double u) {
    const double E = fmax(m0 > 0.0 ? e0 : -__builtin_inf(), fmax(m1 > 0.0 ? e1 : -__builtin_inf(),
                                                                  m2 > 0.0 ? e2 : -__builtin_inf()));
    return select3(xalign(m0, e0, E), xalign(m1, e1, E), xalign(m2, e2, E), u);
}

// Maps of {0, 1, 2} into itself, 2 bits per argument (n -> bits 2n..2n+1);
// compose3(G, H) is n -> G(H(n)).
constexpr int kIdentity3 = 0 | (1 << 2) | (2 << 4);
__device__ __forceinline__ int compose3(int G, int H) {
    int r = 0;
#pragma unroll
    for (int n = 0; n < 3; n++) {
        const int h = (H >> (2 * n)) & 3;
        r |= ((G >> (2 * h)) & 3) << (2 * n);
    }
    return r;
}

// Sums of kFB per-lane values over the wave in 10 exchanges: halving the
// set of values at each of the first three steps; lane l ends with the total
// of value (l >> 3) & 7.
constexpr int kFB = 8;  // frames per wave batch
__device__ __forceinline__ double wave_sum8(const double (&a)[kFB], int lane) {
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8;
    double v4[4], v2[2];
#pragma unroll
    for (int j = 0; j < 4; j++) v4[j] = (b5 ? a[j + 4] : a[j]) + __shfl_xor(b5 ? a[j] : a[j + 4], 32);
#pragma unroll
    for (int j = 0; j < 2; j++) v2[j] = (b4 ? v4[j + 2] : v4[j]) + __shfl_xor(b4 ? v4[j] : v4[j + 2], 16);
    double v = (b3 ? v2[1] : v2[0]) + __shfl_xor(b3 ? v2[0] : v2[1], 8);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 1);
    return v;
}

// ---------------------------------------------------------------------------
// statistics of the plan.  The pixels of a chain are cut into ranges of
// kRange; a list entry is a (spot, range) whose mask has pixels in the
// range, with the segment of the (increasing) mask that falls there.
// b3_gibbs_mask_values: the spot values at the mask pixels and S2;
// b3_gibbs_cpart: per (range, group of kFB frames), each entry's sums;
// b3_gibbs_opart: per entry, its dot products with the other entries of
// the range; b3_gibbs_reduce: C and O per spot, the ranges added in order.

// Workgroups are dispatched to the 8 XCDs round-robin (b mod 8).  Giving
// each XCD a contiguous range of tasks keeps the tasks of one chain, which
// read the same frames and masks, on one L2.  The grid is a multiple of 8;
// returns the task index, or -1 for the padding.
__device__ __forceinline__ int xcd_task(int ntasks) {
    const int per = gridDim.x >> 3, b = blockIdx.x;
    const int t = (b & 7) * per + (b >> 3);
    return t < ntasks ? t : -1;
}

// mu[e] = spot[k][midx[e]] for the entries of mask k (task.k = k), and
// S2_k = sum of their squares.
__global__ __launch_bounds__(kStatThreads) void b3_gibbs_mask_values(const StatDev *chains, const SpotTask *tasks,
                                                                    int ntasks) {
    const int ti = xcd_task(ntasks);
    if (ti < 0) return;
    const SpotTask tk = tasks[ti];
    const StatDev c = chains[tk.chain];
    const int k = tk.k, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ double red[kStatThreads / 64];
    const double *spk = c.spot + (size_t)k * c.P;
    double s2 = 0.0;
    for (int e = c.moff[k] + tid; e < c.moff[k + 1]; e += kStatThreads) {
        const double v = spk[c.midx[e]];
        c.mu[e] = v;
        s2 += v * v;
    }
    s2 = wave_sum(s2);
    if (lane == 0) red[wave] = s2;
    __syncthreads();
    if (tid == 0) {
        double a = 0.0;
        for (int w = 0; w < kStatThreads / 64; w++) a += red[w];
        c.S2[k] = a;
    }
}

// Per (range r, frames t0..t0+kFB-1): the frames over the range are staged
// in LDS with coalesced reads (each pixel value is read from HBM once);
// each wave takes list entries of the range and sums mu (data - base) over
// the entry's segment for every frame, four mask entries per lane in
// flight, into Cp[entry][t].
__global__ __launch_bounds__(kCpartThreads) void b3_gibbs_cpart(const StatDev *chains, const StatTask *tasks,
                                                              int ntasks) {
    __shared__ float fr[kFB * kRange];
    const int ti = xcd_task(ntasks);
    if (ti < 0) return;
    const StatTask tk = tasks[ti];
    const StatDev c = chains[tk.chain];
    const int T = c.T, P = c.P, r = tk.r;
    const int t0 = tk.g * kFB, nf = min(kFB, T - t0);
    const int p0 = r * kRange, np = min(kRange, P - p0);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    {
        // every load of the group in flight at once; rows past T repeat the
        // last frame and are discarded
        constexpr int kU = kRange / kCpartThreads;
        float v[kU][kFB];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int i = tid + u * kCpartThreads;
#pragma unroll
            for (int f = 0; f < kFB; f++) v[u][f] = i < np ? c.data[(size_t)(t0 + min(f, nf - 1)) * P + p0 + i] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kU; u++)
#pragma unroll
            for (int f = 0; f < kFB; f++) fr[f * kRange + tid + u * kCpartThreads] = v[u][f];
    }
    __syncthreads();
    for (int q = c.roff[r] + wave; q < c.roff[r + 1]; q += kCpartThreads / 64) {
        const int e0 = c.lst[4 * q + 2], e1 = c.lst[4 * q + 3];
        double acc[kFB];
#pragma unroll
        for (int f = 0; f < kFB; f++) acc[f] = 0.0;
        for (int e = e0 + lane; e < e1; e += 4 * 64) {
            int x[4];
            double mu[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int eu = e + 64 * u;
                const bool in = eu < e1;
                x[u] = in ? c.midx[eu] - p0 : 0;
                mu[u] = in ? c.mu[eu] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const bool in = e + 64 * u < e1;
#pragma unroll
                for (int f = 0; f < kFB; f++) {
                    const double term = mu[u] * ((double)fr[f * kRange + x[u]] - c.base);
                    acc[f] += in ? term : 0.0;
                }
            }
        }
        const double v = wave_sum8(acc, lane);
        const int f = (lane >> 3) & 7;
        if ((lane & 7) == 0 && f < nf) c.Cp[(size_t)q * T + t0 + f] = v;
    }
}

// Per list entry q = (k, r): mu_k over its segment scattered into a dense
// LDS row (0 elsewhere in the range); each wave takes the entries (j, r) of
// the range with j > k (O is symmetric) and sums mu_j times the row over
// their segments, four mask entries per lane in flight.  Op[q][j] for every
// j (0 for j <= k and for the spots that do not meet the range).
__global__ __launch_bounds__(kOpartThreads) void b3_gibbs_opart(const StatDev *chains, const SpotTask *tasks,
                                                              int ntasks) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int ti = xcd_task(ntasks);
    if (ti < 0) return;
    const SpotTask tk = tasks[ti];
    const StatDev c = chains[tk.chain];
    const int q = tk.k, K = c.K;
    const int k = c.lst[4 * q], r = c.lst[4 * q + 1], ke0 = c.lst[4 * q + 2], ke1 = c.lst[4 * q + 3];
    const int p0 = r * kRange;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    double *row = (double *)lds;  // [kRange]
    double *Ol = row + kRange;    // [K]
    for (int i = tid; i < kRange; i += kOpartThreads) row[i] = 0.0;
    for (int j = tid; j < K; j += kOpartThreads) Ol[j] = 0.0;
    __syncthreads();
    for (int e = ke0 + tid; e < ke1; e += kOpartThreads) row[c.midx[e] - p0] = c.mu[e];
    __syncthreads();
    const int q0 = c.roff[r], nq = c.roff[r + 1] - q0;
    for (int b = 0; b < nq; b += 64) {
        // up to 64 entries of the range, one per lane, then taken in turn
        int lj = 0, l0 = 0, l1 = 0;
        if (b + lane < nq) {
            const int4 ent = *(const int4 *)(c.lst + 4 * (q0 + b + lane));
            lj = ent.x; l0 = ent.z; l1 = ent.w;
        }
        for (int qi = wave; qi < min(64, nq - b); qi += kOpartThreads / 64) {
            const int j = __builtin_amdgcn_readlane(lj, qi), e0 = __builtin_amdgcn_readlane(l0, qi),
                      e1 = __builtin_amdgcn_readlane(l1, qi);
            if (j <= k) continue;  // O is symmetric: the pairs j > k only
            double a = 0.0;
            for (int e = e0 + lane; e < e1; e += 4 * 64) {
                int x[4];
                double mu[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int eu = e + 64 * u;
                    const bool in = eu < e1;
                    x[u] = in ? c.midx[eu] - p0 : 0;
                    mu[u] = in ? c.mu[eu] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const double term = mu[u] * row[x[u]];
                    a += e + 64 * u < e1 ? term : 0.0;
                }
            }
            a = wave_sum(a);
            if (lane == 0) Ol[j] = a;
        }
    }
    __syncthreads();
    for (int j = tid; j < K; j += kOpartThreads) c.Op[(size_t)q * K + j] = Ol[j];
}

// C[k][t] and O[k][j] (task.k = k): the partial sums of spot k's list
// entries added in range order.
__global__ __launch_bounds__(kStatThreads) void b3_gibbs_reduce(const StatDev *chains, const SpotTask *tasks,
                                                               int ntasks) {
    const int ti = xcd_task(ntasks);
    if (ti < 0) return;
    const SpotTask tk = tasks[ti];
    const StatDev c = chains[tk.chain];
    const int k = tk.k, T = c.T, K = c.K;
    const int q0 = c.koff[k], q1 = c.koff[k + 1];
    for (int t = threadIdx.x; t < T; t += kStatThreads) {
        double a = 0.0;
        for (int q = q0; q < q1; q++) a += c.Cp[(size_t)c.kent[q] * T + t];
        c.C[(size_t)k * T + t] = a;
    }
    // O_kj = O_jk: this spot writes both for j > k (the partial sums hold the
    // pairs j > k only), and 0 on the diagonal
    for (int j = k + threadIdx.x; j < K; j += kStatThreads) {
        double a = 0.0;
        if (j > k)
            for (int q = q0; q < q1; q++) a += c.Op[(size_t)c.kent[q] * K + j];
        c.O[(size_t)k * K + j] = a;
        c.O[(size_t)j * K + k] = a;
    }
}

// ---------------------------------------------------------------------------
// the sweeps: one 64-thread workgroup per chain
//
// Per spot update:
// 1. forward messages, frames taken 128 at a time, two per lane: each lane
//    forms its frames' x = B_on - B_off, b_on = e^min(x,0), b_off =
//    e^min(-x,0) and transfer matrices M_t[j][i] = b_j A[i][j] (b_j pi_j in
//    frame 0), in extended range, and their product; a 6-step scan across
//    the wave forms the prefix products, and the message of the frame
//    before the group carried through the products of the lanes before and
//    the lane's own two matrices gives its frames' forward messages
//    (forward_algorithm_delta2's delta_t, forward_algorithm.h:344-382, up to
//    a per-frame factor, which leaves sample_unscaled_log's distribution
//    unchanged);
// 2. backward sampling, frames taken 64 at a time, one per lane
//    (forward_algorithm.h:542-566): the last frame is drawn
//    from alpha_{T-1} and frame t from alpha_t(i) A[i][s_{t+1}] with draw
//    u_t; each lane evaluates select_random_element for its frame under
//    each of the three possible s_{t+1} (the arithmetic of the sequential
//    rule), and a 6-step scan composes these maps, so that every frame's
//    state follows from the state after the group;
// 3. the frames whose on/off status changed are compacted with a ballot,
//    and A of the other spots is updated there.

__global__ __launch_bounds__(kSweepThreads) void b3_gibbs(GibbsParams G) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x;
    const GibbsDev c = G.chains[G.order[blockIdx.x]];
    const int T = c.T, K = c.K, Tc = G.Tcap;
    const int RC = Tc + kMtN;  // draw ring capacity

    double *am = (double *)lds;             // [Tc][3] forward messages, extended range:
    double *ae = am + 3 * Tc;               // [Tc][3]   am 2^ae
    int *chg = (int *)(ae + 3 * Tc);        // [Tc] frames whose on/off status changed
    uint32_t *mt = (uint32_t *)(chg + Tc);  // [kMtN]
    uint32_t *ring = mt + kMtN;             // [RC]
    const double *A = G.A, *pi = G.pi;
    double *__restrict__ Aall = c.A;
    const double *__restrict__ Oall = c.O;

    for (int i = lane; i < kMtN; i += kSweepThreads) mt[i] = c.mt[i];
    const int pos0 = (int)c.mt[kMtN];
    __syncthreads();
    // Every thread tracks the ring's head and count; they change uniformly.
    int head = 0, cnt = kMtN - pos0;
    for (int i = lane; i < cnt; i += kSweepThreads) ring[i] = mt_temper(mt[pos0 + i]);
    const double scale = 1.0 / (65536.0 * 65536.0);
    const int tlast = ((T - 1) / kSweepThreads) * kSweepThreads;

    for (int s = 0; s < G.samples; s++) {
        const double i2v = 0.5 / G.var[s];
        for (int it = 0; it < G.iters; it++) {
            for (int k = 0; k < K; k++) {
                const double *Ck = c.C + (size_t)k * T;
                const double *Ak = Aall + (size_t)k * T;
                const double S2 = c.S2[k];
                uint8_t *smp = c.sample + (size_t)k * T;
                // T draws in the ring
                while (cnt < T) {
                    mt_twist64(mt, lane);
                    for (int i = lane; i < kMtN; i += kSweepThreads) {
                        int w = head + cnt + i;
                        w = w >= RC ? w - RC : w;
                        ring[w] = mt_temper(mt[i]);
                    }
                    cnt += kMtN;
                }
                // 1. forward messages: per group of 128 frames, two per lane:
                //    the lane's product of its two transfer matrices, prefix
                //    products of those across the wave, and the message of
                //    the frame before the group (the carry; (1, 0, 0) for the
                //    first group, whose first matrix ignores it) carried
                //    through the lane's two frames
                const double ninf = -__builtin_inf();
                double cm[3] = {0.5, 0.0, 0.0}, ce[3] = {1.0, ninf, ninf};
                for (int t0 = 0; t0 < T; t0 += 2 * kSweepThreads) {
                    const int ta = t0 + 2 * lane, tb = ta + 1;
                    double Pm[9], Pe[9];
                    {
                        double Mam[9], Mae[9], Mbm[9], Mbe[9];
                        frame_matrix(ta, T, S2, Ck, Ak, i2v, A, pi, Mam, Mae);
                        frame_matrix(tb, T, S2, Ck, Ak, i2v, A, pi, Mbm, Mbe);
                        xmatmul(Mbm, Mbe, Mam, Mae, Pm, Pe);
                    }
                    prefix_products<true>(Pm, Pe, lane);
                    // the product over the lanes before this one (identity for lane 0)
                    double Em[9], Ee[9];
#pragma unroll
                    for (int q = 0; q < 9; q++) {
                        const double m = __shfl_up(Pm[q], 1), e = __shfl_up(Pe[q], 1);
                        Em[q] = lane == 0 ? (q % 4 == 0 ? 0.5 : 0.0) : m;
                        Ee[q] = lane == 0 ? (q % 4 == 0 ? 1.0 : ninf) : e;
                    }
                    double vm[3], ve[3], wam[3], wae[3], wbm[3], wbe[3];
                    xmatvec(Em, Ee, cm, ce, vm, ve);
                    {
                        double Mm[9], Me[9];
                        frame_matrix(ta, T, S2, Ck, Ak, i2v, A, pi, Mm, Me);
                        xmatvec(Mm, Me, vm, ve, wam, wae);
                        frame_matrix(tb, T, S2, Ck, Ak, i2v, A, pi, Mm, Me);
                        xmatvec(Mm, Me, wam, wae, wbm, wbe);
                    }
                    if (ta < T) {
#pragma unroll
                        for (int j = 0; j < 3; j++) {
                            am[3 * ta + j] = wam[j];
                            ae[3 * ta + j] = wae[j];
                        }
                    }
                    if (tb < T) {
#pragma unroll
                        for (int j = 0; j < 3; j++) {
                            am[3 * tb + j] = wbm[j];
                            ae[3 * tb + j] = wbe[j];
                        }
                    }
                    // carry: the message of the group's last frame
                    const int tl = min(T, t0 + 2 * kSweepThreads) - 1 - t0;
                    const int ll = tl >> 1;
                    const bool odd = tl & 1;
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        cm[j] = odd ? read_lane(wbm[j], ll) : read_lane(wam[j], ll);
                        ce[j] = odd ? read_lane(wbe[j], ll) : read_lane(wae[j], ll);
                    }
                }
                __syncthreads();
                // 2. backward sampling, last frame first; frame t takes draw
                //    head + (T - 1 - t)
                int sn = 0, nchg = 0;
                for (int t0 = tlast; t0 >= 0; t0 -= kSweepThreads) {
                    const int t = t0 + lane;
                    int f = 0;  // the state drawn as a map of s_{t+1}
                    if (t < T) {
                        int r = head + (T - 1 - t);
                        r = r >= RC ? r - RC : r;
                        const double u = (double)ring[r] * scale;
                        const double v0 = am[3 * t], v1 = am[3 * t + 1], v2 = am[3 * t + 2];
                        const double e0 = ae[3 * t], e1 = ae[3 * t + 1], e2 = ae[3 * t + 2];
                        if (t == T - 1) {
                            f = xselect3(v0, e0, v1, e1, v2, e2, u) * 21;
                        } else {
#pragma unroll
                            for (int n = 0; n < 3; n++)
                                f |= xselect3(v0 * A[n], e0, v1 * A[3 + n], e1, v2 * A[6 + n], e2, u) << (2 * n);
                        }
                    }
                    // suffix compositions f_t o f_{t+1} o ... over the group,
                    // applied to the state of the frame after the group
                    int G = t < T ? f : kIdentity3;
#pragma unroll
                    for (int d = 1; d < kSweepThreads; d <<= 1) {
                        const int H = __shfl_down(G, d);
                        if (lane + d < kSweepThreads) G = compose3(G, H);
                    }
                    const int sv = (G >> (2 * sn)) & 3;
                    sn = __builtin_amdgcn_readlane(sv, 0);
                    bool ch = false;
                    int code = 0;
                    if (t < T) {
                        const int old = smp[t];
                        smp[t] = (uint8_t)sv;
                        ch = (old == 0) != (sv == 0);
                        code = (t << 1) | (old == 0 ? 1 : 0);
                    }
                    const unsigned long long bal = __ballot(ch);
                    if (ch) chg[nchg + __popcll(bal & ((1ull << lane) - 1))] = code;
                    nchg += __popcll(bal);
                }
                head += T;
                head = head >= RC ? head - RC : head;
                cnt -= T;
                __syncthreads();
                // 3. A of the other spots in the frames where this spot
                //    turned on (+O) or off (-O)
                for (int e = lane; e < nchg * K; e += kSweepThreads) {
                    const int f = e / K, j = e - f * K;
                    if (j == k) continue;
                    const int code = chg[f];
                    const double o = Oall[(size_t)j * K + k];
                    double *a = Aall + (size_t)j * T + (code >> 1);
                    *a = (code & 1) ? *a - o : *a + o;
                }
                __syncthreads();
            }
        }
        if (c.samples) {
            uint8_t *dst = c.samples + (size_t)s * K * T;
            for (int i = lane; i < K * T; i += kSweepThreads) dst[i] = c.sample[i];
            __syncthreads();
        }
    }
    // the generator's state and the position of its next draw
    for (int i = lane; i < kMtN; i += kSweepThreads) c.mt[i] = mt[i];
    if (lane == 0) c.mt[kMtN] = (uint32_t)(kMtN - cnt);
}

// GibbsSampler2::reset (sampled_multispot.h:430-438): all spots in state 2,
// no spot in the sample intensities (A = 0).
__global__ void b3_gibbs_reset(const GibbsDev *chains) {
    const GibbsDev c = chains[blockIdx.y];
    const size_t n = (size_t)c.K * c.T;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        c.A[e] = 0.0;
        c.sample[e] = 2;
    }
}

// Sample intensities of chain i from its sample: base, plus each on spot
// over its mask, spots in index order (a barrier between spots: masks
// overlap).
__global__ __launch_bounds__(kStatThreads) void b3_gibbs_export(const GibbsDev *chains, int i, double *out) {
    const GibbsDev c = chains[i];
    const int T = c.T, P = c.P;
    for (size_t e = threadIdx.x; e < (size_t)T * P; e += kStatThreads) out[e] = c.base;
    __syncthreads();
    for (int j = 0; j < c.K; j++) {
        const int m0 = c.moff[j], m1 = c.moff[j + 1];
        const double *spj = c.spot + (size_t)j * P;
        for (int t = 0; t < T; t++) {
            if (c.sample[(size_t)j * T + t] != 0) continue;
            double *o = out + (size_t)t * P;
            for (int q = m0 + threadIdx.x; q < m1; q += kStatThreads) {
                const int p = c.midx[q];
                o[p] += spj[p];
            }
        }
        __syncthreads();
    }
}

// compute_spot_intensity: spot_shape (storm.h:51-57, 229-235) at integer
// pixel coordinates, FP64.
__global__ void b3_spot_intensity(const int *__restrict__ xyz, int P, const double *__restrict__ phi, int K,
                                  double zF, double *__restrict__ out) {
    const size_t n = (size_t)P * K, stride = (size_t)gridDim.x * blockDim.x;
    const double r_2_pi = sqrt(2.0 * M_PI);
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        const int k = (int)(e / P), p = (int)(e - (size_t)k * P);
        const double *ph = phi + 5 * (size_t)k;
        const int *x = xyz + 3 * (size_t)p;
        const double d0 = x[0] - ph[2], d1 = x[1] - ph[3];
        double d2 = x[2] - ph[4];
        d2 = d2 / zF;
        const double s = -(d0 * d0 + d1 * d1 + d2 * d2) / (2 * ph[1] * ph[1]);
        out[e] = exp(s) * ph[0] / (ph[1] * r_2_pi);
    }
}

}  // namespace rfba3b

// ---------------------------------------------------------------------------
// host side

using namespace rfba3b;

#define GHIP(expr)                                                                                       \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return host_fail(e_ == hipErrorOutOfMemory ? -2 : -3, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace {

// Host MT19937 as the reference seeds it (mersenne.cpp Init0 + RandomInit's
// 37 discarded draws), then `skip` further draws; the state is handed to the
// device as 624 words plus the position of the next draw.
struct HostMt {
    uint32_t mt[kMtN];
    int pos;
    explicit HostMt(uint32_t seed) {
        mt[0] = seed;
        for (int i = 1; i < kMtN; i++) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
        pos = kMtN;
    }
    void twist() {
        auto word = [](uint32_t cur, uint32_t nxt, uint32_t far) {
            const uint32_t y = (cur & 0x80000000u) | (nxt & 0x7fffffffu);
            return far ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        };
        int kk = 0;
        for (; kk < kMtN - 397; kk++) mt[kk] = word(mt[kk], mt[kk + 1], mt[kk + 397]);
        for (; kk < kMtN - 1; kk++) mt[kk] = word(mt[kk], mt[kk + 1], mt[kk + 397 - kMtN]);
        mt[kMtN - 1] = word(mt[kMtN - 1], mt[0], mt[396]);
        pos = 0;
    }
    void discard(unsigned long long n) {
        while (n > 0) {
            if (pos >= kMtN) twist();
            const unsigned long long take = std::min<unsigned long long>(n, (unsigned long long)(kMtN - pos));
            pos += (int)take;
            n -= take;
        }
    }
};

// LDS of the sweep kernel for frames up to Tc
size_t gibbs_lds(int Tc) { return 16 * (3 * (size_t)Tc) + 4 * (size_t)Tc + 4 * ((size_t)kMtN + Tc + kMtN); }

// LDS of the overlap kernel for up to K spots per chain
size_t overlap_lds(int K) { return 8 * ((size_t)kRange + K); }

}  // namespace

struct rfba3b_gibbs {
    rfba3b_model model;
    std::vector<rfba3b_gibbs_chain> chains;
    std::vector<unsigned long long> draws;
    GibbsDev *d_chains = nullptr;
    int *d_order = nullptr;
    uint32_t *d_mt = nullptr;
    int *d_ints = nullptr;       // every chain's mask offsets and indices
    double *d_dbl = nullptr;     // every chain's C, A, O, S2
    double *d_var = nullptr;
    int var_cap = 0;
    int Tmax = 0;
    size_t lds = 0;
    ~rfba3b_gibbs() {
        if (d_chains) (void)hipFree(d_chains);
        if (d_order) (void)hipFree(d_order);
        if (d_mt) (void)hipFree(d_mt);
        if (d_ints) (void)hipFree(d_ints);
        if (d_dbl) (void)hipFree(d_dbl);
        if (d_var) (void)hipFree(d_var);
    }
};

static int gibbs_create(const rfba3b_gibbs_chain *chains, int n, const rfba3b_model *model, rfba3b_gibbs **out,
                        float *setup_ms) {
    if (!out) return host_fail(-1, "plan pointer is NULL");
    *out = nullptr;
    if (setup_ms) *setup_ms = 0.0f;
    if (!chains || n <= 0) return host_fail(-1, "no chains");
    if (n > 65535) return host_fail(-1, "%d chains: at most 65535 per plan", n);
    if (int rc = host_check_model(model)) return rc;
    int ncu = 0;
    if (int rc = host_dev_init(&ncu)) return rc;
    int Tmax = 1, Kmax = 0;
    for (int i = 0; i < n; i++) {
        const rfba3b_gibbs_chain &c = chains[i];
        if (c.P <= 0 || c.T <= 0 || c.K < 0) return host_fail(-1, "chain %d: P and T must be positive, K >= 0", i);
        if ((size_t)c.P * c.T > (size_t)INT_MAX * 64) return host_fail(-1, "chain %d: T x P too large", i);
        if (!c.d_data || !c.d_sample) return host_fail(-1, "chain %d: data or sample is NULL", i);
        if (c.K > 0 && (!c.d_spot || !c.d_mask_off || !c.d_mask))
            return host_fail(-1, "chain %d: spot table or masks are NULL", i);
        if (!std::isfinite(c.base)) return host_fail(-1, "chain %d: base is not finite", i);
        Tmax = std::max(Tmax, c.T);
        Kmax = std::max(Kmax, c.K);
    }
    const size_t lds = gibbs_lds(Tmax), olds = overlap_lds(Kmax);
    if (lds > 160 * 1024)
        return host_fail(-1, "T = %d needs %zu bytes of LDS per chain (at most 160 KiB)", Tmax, lds);
    if (olds > 160 * 1024)
        return host_fail(-1, "K = %d needs %zu bytes of LDS in the overlaps (at most 160 KiB)", Kmax, olds);

    auto *g = new rfba3b_gibbs();
    std::unique_ptr<rfba3b_gibbs> own(g);
    g->model = *model;
    g->chains.assign(chains, chains + n);
    g->draws.resize(n);
    g->Tmax = Tmax;
    g->lds = lds;

    // masks to the host for checking; offsets and indices kept by the plan,
    // with the statistics' range lists
    std::vector<int> ints;
    std::vector<size_t> o_off(n), o_idx(n), o_roff(n), o_lst(n), o_koff(n), o_kent(n), o_dbl(n), o_scr(n);
    std::vector<int> nent(n), nmask(n);
    std::vector<long long> work(n);
    std::vector<SpotTask> stasks, etasks;  // (chain, spot), (chain, list entry)
    std::vector<StatTask> ctasks;          // (chain, range, frame group)
    size_t ndbl = 0, nscr = 0;
    for (int i = 0; i < n; i++) {
        const rfba3b_gibbs_chain &c = chains[i];
        std::vector<int> off(c.K + 1, 0), mask;
        if (c.K > 0) {
            GHIP(hipMemcpy(off.data(), c.d_mask_off, (c.K + 1) * sizeof(int), hipMemcpyDeviceToHost));
            if (off[0] != 0) return host_fail(-1, "chain %d: mask offsets must start at 0", i);
            for (int k = 0; k < c.K; k++)
                if (off[k + 1] < off[k]) return host_fail(-1, "chain %d: mask offsets decrease at spot %d", i, k);
            mask.resize(off[c.K]);
            if (!mask.empty())
                GHIP(hipMemcpy(mask.data(), c.d_mask, mask.size() * sizeof(int), hipMemcpyDeviceToHost));
            for (int k = 0; k < c.K; k++)
                for (int j = off[k]; j < off[k + 1]; j++)
                    if (mask[j] < 0 || mask[j] >= c.P || (j > off[k] && mask[j - 1] >= mask[j]))
                        return host_fail(-1, "chain %d: mask of spot %d is not increasing within [0, P) at %d", i,
                                         k, j - off[k]);
        }
        o_off[i] = ints.size();
        ints.insert(ints.end(), off.begin(), off.end());
        o_idx[i] = ints.size();
        ints.insert(ints.end(), mask.begin(), mask.end());
        // list entries (spot, range, segment) by range, then spot
        const int R = (c.P + kRange - 1) / kRange;
        std::vector<int> roff(R + 1, 0), lst, koff(c.K + 1, 0), kent;
        std::vector<std::vector<int>> per_spot(c.K);
        std::vector<int> cur(off.begin(), off.end() - 1);  // next mask entry of each spot
        for (int r = 0; r < R; r++) {
            roff[r] = (int)(lst.size() / 4);
            const int pend = std::min((r + 1) * kRange, c.P);
            for (int k = 0; k < c.K; k++) {
                int e = cur[k];
                while (e < off[k + 1] && mask[e] < pend) e++;
                if (e > cur[k]) {
                    per_spot[k].push_back((int)(lst.size() / 4));
                    lst.insert(lst.end(), {k, r, cur[k], e});
                }
                cur[k] = e;
            }
        }
        roff[R] = (int)(lst.size() / 4);
        for (int k = 0; k < c.K; k++) {
            koff[k] = (int)kent.size();
            kent.insert(kent.end(), per_spot[k].begin(), per_spot[k].end());
        }
        koff[c.K] = (int)kent.size();
        nent[i] = roff[R];
        nmask[i] = off[c.K];
        o_roff[i] = ints.size();
        ints.insert(ints.end(), roff.begin(), roff.end());
        while (ints.size() % 4) ints.push_back(0);  // entries read as int4
        o_lst[i] = ints.size();
        ints.insert(ints.end(), lst.begin(), lst.end());
        o_koff[i] = ints.size();
        ints.insert(ints.end(), koff.begin(), koff.end());
        o_kent[i] = ints.size();
        ints.insert(ints.end(), kent.begin(), kent.end());
        o_dbl[i] = ndbl;  // C [K][T], A [K][T], O [K][K], S2 [K], 256-byte aligned
        ndbl += ((2 * (size_t)c.K * c.T + (size_t)c.K * c.K + c.K) + 31) & ~(size_t)31;
        o_scr[i] = nscr;  // mu [mask entries], Cp [ne][T], Op [ne][K]
        nscr += ((size_t)off[c.K] + (size_t)nent[i] * (c.T + c.K) + 31) & ~(size_t)31;
        for (int k = 0; k < c.K; k++) stasks.push_back({i, k});
        for (int q = 0; q < nent[i]; q++) etasks.push_back({i, q});
        for (int r = 0; r < R; r++)
            if (roff[r + 1] > roff[r])
                for (int b = 0; b < (c.T + kFB - 1) / kFB; b++) ctasks.push_back({i, r, b});
        work[i] = (long long)c.K * c.T;
    }
    GHIP(hipMalloc(&g->d_ints, std::max<size_t>(ints.size(), 1) * sizeof(int)));
    if (!ints.empty()) GHIP(hipMemcpy(g->d_ints, ints.data(), ints.size() * sizeof(int), hipMemcpyHostToDevice));
    GHIP(hipMalloc(&g->d_dbl, std::max<size_t>(ndbl, 1) * sizeof(double)));

    std::vector<uint32_t> mts((size_t)n * (kMtN + 1));
    std::vector<GibbsDev> dev(n);
    GHIP(hipMalloc(&g->d_mt, mts.size() * sizeof(uint32_t)));
    for (int i = 0; i < n; i++) {
        const rfba3b_gibbs_chain &c = chains[i];
        HostMt h(c.seed);
        h.discard(37ull + c.skip);
        memcpy(&mts[(size_t)i * (kMtN + 1)], h.mt, sizeof h.mt);
        mts[(size_t)i * (kMtN + 1) + kMtN] = (uint32_t)h.pos;
        g->draws[i] = c.skip;
        GibbsDev &d = dev[i];
        d.P = c.P; d.T = c.T; d.K = c.K;
        d.data = c.d_data; d.spot = c.d_spot;
        d.moff = g->d_ints + o_off[i];
        d.midx = g->d_ints + o_idx[i];
        d.sample = c.d_sample; d.samples = c.d_samples;
        double *b = g->d_dbl + o_dbl[i];
        d.C = b;
        d.A = b + (size_t)c.K * c.T;
        d.O = b + 2 * (size_t)c.K * c.T;
        d.S2 = d.O + (size_t)c.K * c.K;
        d.base = c.base;
        d.mt = g->d_mt + (size_t)i * (kMtN + 1);
    }
    GHIP(hipMemcpy(g->d_mt, mts.data(), mts.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    GHIP(hipMalloc(&g->d_chains, n * sizeof(GibbsDev)));
    GHIP(hipMemcpy(g->d_chains, dev.data(), n * sizeof(GibbsDev), hipMemcpyHostToDevice));

    // statistics, in buffers that live until they are done
    if (!stasks.empty()) {
        struct DevFree {
            void operator()(void *p) const { (void)hipFree(p); }
        };
        std::vector<std::unique_ptr<void, DevFree>> tmp;
        auto dev_alloc = [&](size_t bytes) -> void * {
            void *p = nullptr;
            if (hipMalloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) return nullptr;
            tmp.emplace_back(p);
            return p;
        };
        double *scr = (double *)dev_alloc(nscr * sizeof(double));
        StatDev *d_sdev = (StatDev *)dev_alloc(n * sizeof(StatDev));
        SpotTask *d_st = (SpotTask *)dev_alloc(stasks.size() * sizeof(SpotTask));
        SpotTask *d_et = (SpotTask *)dev_alloc(etasks.size() * sizeof(SpotTask));
        StatTask *d_ct = (StatTask *)dev_alloc(ctasks.size() * sizeof(StatTask));
        if (!scr || !d_sdev || !d_st || !d_et || !d_ct) return host_fail(-2, "out of device memory (statistics)");
        std::vector<StatDev> sdev(n);
        for (int i = 0; i < n; i++) {
            const rfba3b_gibbs_chain &c = chains[i];
            StatDev &s = sdev[i];
            s.P = c.P; s.T = c.T; s.K = c.K; s.ne = nent[i];
            s.data = c.d_data; s.spot = c.d_spot;
            s.moff = dev[i].moff; s.midx = dev[i].midx;
            s.base = c.base;
            s.roff = g->d_ints + o_roff[i];
            s.lst = g->d_ints + o_lst[i];
            s.koff = g->d_ints + o_koff[i];
            s.kent = g->d_ints + o_kent[i];
            s.mu = scr + o_scr[i];
            s.Cp = s.mu + nmask[i];
            s.Op = s.Cp + (size_t)s.ne * c.T;
            s.C = dev[i].C; s.O = dev[i].O; s.S2 = dev[i].S2;
        }
        GHIP(hipMemcpy(d_sdev, sdev.data(), n * sizeof(StatDev), hipMemcpyHostToDevice));
        GHIP(hipMemcpy(d_st, stasks.data(), stasks.size() * sizeof(SpotTask), hipMemcpyHostToDevice));
        if (!etasks.empty())
            GHIP(hipMemcpy(d_et, etasks.data(), etasks.size() * sizeof(SpotTask), hipMemcpyHostToDevice));
        if (!ctasks.empty())
            GHIP(hipMemcpy(d_ct, ctasks.data(), ctasks.size() * sizeof(StatTask), hipMemcpyHostToDevice));
        if (olds > 64 * 1024)
            GHIP(hipFuncSetAttribute((const void *)b3_gibbs_opart, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)olds));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
        if (e == hipSuccess) {
            auto grid = [](size_t nt) { return dim3((unsigned)((nt + 7) / 8 * 8)); };
            const int ns = (int)stasks.size(), ne = (int)etasks.size(), nc = (int)ctasks.size();
            hipLaunchKernelGGL(b3_gibbs_mask_values, grid(ns), dim3(kStatThreads), 0, nullptr,
                               (const StatDev *)d_sdev, (const SpotTask *)d_st, ns);
            if (nc > 0)
                hipLaunchKernelGGL(b3_gibbs_cpart, grid(nc), dim3(kCpartThreads), 0, nullptr, (const StatDev *)d_sdev,
                                   (const StatTask *)d_ct, nc);
            if (ne > 0)
                hipLaunchKernelGGL(b3_gibbs_opart, grid(ne), dim3(kOpartThreads), olds, nullptr,
                                   (const StatDev *)d_sdev, (const SpotTask *)d_et, ne);
            hipLaunchKernelGGL(b3_gibbs_reduce, grid(ns), dim3(kStatThreads), 0, nullptr, (const StatDev *)d_sdev,
                               (const SpotTask *)d_st, ns);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        GHIP(e);
        if (setup_ms) *setup_ms = ms;
    }

    // launch order: largest work (spots x frames) first
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
    GHIP(hipMalloc(&g->d_order, n * sizeof(int)));
    GHIP(hipMemcpy(g->d_order, order.data(), n * sizeof(int), hipMemcpyHostToDevice));

    if (lds > 64 * 1024)
        GHIP(hipFuncSetAttribute((const void *)b3_gibbs, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    GHIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)b3_gibbs, kSweepThreads, lds));
    if (per_cu <= 0) return host_fail(-1, "Gibbs kernel does not fit on a CU (LDS %zu bytes)", lds);
    hipLaunchKernelGGL(b3_gibbs_reset, dim3(16, n), dim3(256), 0, nullptr, (const GibbsDev *)g->d_chains);
    GHIP(hipGetLastError());
    GHIP(hipDeviceSynchronize());
    *out = own.release();
    return 0;
}

extern "C" int rfba3b_gibbs_create(const rfba3b_gibbs_chain *chains, int n, const rfba3b_model *model,
                                   rfba3b_gibbs **plan, float *setup_ms) {
    std::lock_guard<std::mutex> lk(host_mutex());
    return gibbs_create(chains, n, model, plan, setup_ms);
}

extern "C" int rfba3b_gibbs_reset(rfba3b_gibbs *g, void *stream) {
    std::lock_guard<std::mutex> lk(host_mutex());
    if (!g) return host_fail(-1, "plan is NULL");
    const int n = (int)g->chains.size();
    hipLaunchKernelGGL(b3_gibbs_reset, dim3(16, n), dim3(256), 0, (hipStream_t)stream,
                       (const GibbsDev *)g->d_chains);
    GHIP(hipGetLastError());
    return 0;
}

extern "C" int rfba3b_gibbs_run(rfba3b_gibbs *g, int samples, int sample_iterations, const double *variance,
                                void *stream, float *kernel_ms) {
    std::lock_guard<std::mutex> lk(host_mutex());
    if (!g) return host_fail(-1, "plan is NULL");
    if (samples < 0 || sample_iterations < 0) return host_fail(-1, "samples and sample_iterations must be >= 0");
    if (kernel_ms) *kernel_ms = 0.0f;
    if (samples == 0) return 0;
    std::vector<double> var(samples, g->model.variance);
    if (variance)
        for (int s = 0; s < samples; s++) {
            if (!(variance[s] > 0) || !std::isfinite(variance[s]))
                return host_fail(-1, "variance[%d] must be positive and finite", s);
            var[s] = variance[s];
        }
    hipStream_t st = (hipStream_t)stream;
    if (samples > g->var_cap) {
        if (g->d_var) GHIP(hipFree(g->d_var));
        g->d_var = nullptr;
        g->var_cap = 0;
        GHIP(hipMalloc(&g->d_var, samples * sizeof(double)));
        g->var_cap = samples;
    }
    GHIP(hipMemcpyAsync(g->d_var, var.data(), samples * sizeof(double), hipMemcpyHostToDevice, st));
    GibbsParams P;
    memset(&P, 0, sizeof P);
    P.chains = g->d_chains;
    P.order = g->d_order;
    P.samples = samples;
    P.iters = sample_iterations;
    P.Tcap = g->Tmax;
    P.var = g->d_var;
    for (int i = 0; i < 9; i++) P.A[i] = g->model.A[i];
    for (int i = 0; i < 3; i++) P.pi[i] = g->model.pi[i];
    hipEvent_t e0, e1;
    GHIP(hipEventCreate(&e0));
    GHIP(hipEventCreate(&e1));
    GHIP(hipEventRecord(e0, st));
    hipLaunchKernelGGL(b3_gibbs, dim3((unsigned)g->chains.size()), dim3(kSweepThreads), g->lds, st, P);
    const hipError_t le = hipGetLastError();
    GHIP(hipEventRecord(e1, st));
    const hipError_t se = hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    GHIP(le);
    GHIP(se);
    for (size_t i = 0; i < g->chains.size(); i++)
        g->draws[i] += (unsigned long long)samples * sample_iterations * g->chains[i].K * g->chains[i].T;
    if (kernel_ms) *kernel_ms = ms;
    return 0;
}

extern "C" int rfba3b_gibbs_sample_intensities(const rfba3b_gibbs *g, int chain, double *d_out, void *stream) {
    std::lock_guard<std::mutex> lk(host_mutex());
    if (!g || !d_out) return host_fail(-1, "plan or output is NULL");
    if (chain < 0 || chain >= (int)g->chains.size()) return host_fail(-1, "chain %d out of range", chain);
    hipLaunchKernelGGL(b3_gibbs_export, dim3(1), dim3(kStatThreads), 0, (hipStream_t)stream,
                       (const GibbsDev *)g->d_chains, chain, d_out);
    GHIP(hipGetLastError());
    return 0;
}

extern "C" int rfba3b_gibbs_draws(const rfba3b_gibbs *g, int chain, unsigned long long *draws) {
    if (!g || !draws) return host_fail(-1, "plan or output is NULL");
    if (chain < 0 || chain >= (int)g->chains.size()) return host_fail(-1, "chain %d out of range", chain);
    *draws = g->draws[chain];
    return 0;
}

extern "C" void rfba3b_gibbs_destroy(rfba3b_gibbs *g) {
    std::lock_guard<std::mutex> lk(host_mutex());
    delete g;
}

// get_spot_pixels (multispot5.cc:1001-1020) over getBall (:1256-1269) and
// ir_rounded3D (:989-995).  The ball radius is a float, as getBall's
// parameter, and its test compares an unsigned squared length converted to
// float with the float square of the radius (image3d.h:78-82).
extern "C" int rfba3b_spot_pixels(const int *xyz, int P, const double *phi, int K, int *offsets, int *idx,
                                  size_t cap) {
    if (P < 0 || K < 0) return host_fail(-1, "P and K must be >= 0");
    if (!offsets || (P > 0 && !xyz) || (K > 0 && !phi)) return host_fail(-1, "NULL buffer");
    auto less = [](const int *a, const int *b) {  // Ref3D::operator< (image3d.h:50-53)
        return a[2] < b[2] || (a[2] == b[2] && (a[1] < b[1] || (a[1] == b[1] && a[0] < b[0])));
    };
    for (int p = 1; p < P; p++)
        if (!less(xyz + 3 * (p - 1), xyz + 3 * p))
            return host_fail(-1, "pixel list is not strictly increasing in (z, y, x) at %d", p);
    // A sorted list of P distinct pixels that fills its bounding box is the
    // box: then a pixel's index follows from its coordinates.
    int b0[3] = {0, 0, 0}, b1[3] = {-1, -1, -1};
    bool box = false;
    if (P > 0) {
        for (int j = 0; j < 3; j++) b0[j] = b1[j] = xyz[j];
        for (int p = 1; p < P; p++)
            for (int j = 0; j < 3; j++) {
                b0[j] = std::min(b0[j], xyz[3 * (size_t)p + j]);
                b1[j] = std::max(b1[j], xyz[3 * (size_t)p + j]);
            }
        box = (long long)(b1[0] - b0[0] + 1) * (b1[1] - b0[1] + 1) * (b1[2] - b0[2] + 1) == (long long)P;
    }
    const long long nx = b1[0] - b0[0] + 1, ny = b1[1] - b0[1] + 1;
    size_t tot = 0;
    offsets[0] = 0;
    for (int k = 0; k < K; k++) {
        const double *ph = phi + 5 * (size_t)k;
        for (int j = 1; j < 5; j++)
            if (!std::isfinite(ph[j])) return host_fail(-1, "spot %d has a non-finite width or position", k);
        const float radius = (float)(ph[1] * 6 + 1);
        const float r2 = radius * radius;
        const float rr = std::ceil(radius + 1);
        if (!(rr <= 4096.0f)) return host_fail(-1, "spot %d: mask radius %g too large", k, (double)radius);
        const int r = (int)rr;
        int off[3];
        for (int j = 0; j < 3; j++) {
            const double v = ph[2 + j];
            off[j] = static_cast<int>(v > 0.0 ? v + 0.5 : v - 0.5);
        }
        // Ball points in (z, y, x) order, as get_ball lists them; the points
        // of one (z, y) row are a contiguous run of x (the test is monotone
        // in |x|), and the pixel list is sorted in (z, y, x), so each row
        // takes one lower_bound and then a forward walk along x.
        for (int z = -r; z <= r; z++)
            for (int y = -r; y <= r; y++) {
                // the largest |x| <= r with (float)(x^2 + y^2 + z^2) <= r2:
                // an integer square root, then the exact test moves it
                const unsigned int c = unsigned(y * y) + unsigned(z * z);
                if (!((float)c <= r2)) continue;
                int w = (int)std::sqrt(std::max(0.0, (double)r2 - (double)c));
                while (w < r && (float)(unsigned(w + 1) * unsigned(w + 1) + c) <= r2) w++;
                while (w > 0 && !((float)(unsigned(w) * unsigned(w) + c) <= r2)) w--;
                w = std::min(w, r);
                const int xl = -w, xh = w;
                if (box) {
                    const int zz = z + off[2], yy = y + off[1];
                    if (zz < b0[2] || zz > b1[2] || yy < b0[1] || yy > b1[1]) continue;
                    const int xa = std::max(xl + off[0], b0[0]), xb = std::min(xh + off[0], b1[0]);
                    const long long row = ((long long)(zz - b0[2]) * ny + (yy - b0[1])) * nx - b0[0];
                    for (int xx = xa; xx <= xb; xx++) {
                        if (idx && tot < cap) idx[tot] = (int)(row + xx);
                        tot++;
                    }
                    continue;
                }
                const int q0[3] = {xl + off[0], y + off[1], z + off[2]};
                int lo = 0, hi = P;  // lower_bound of the row's first point
                while (lo < hi) {
                    const int mid = lo + (hi - lo) / 2;
                    if (less(xyz + 3 * (size_t)mid, q0)) lo = mid + 1;
                    else hi = mid;
                }
                for (int x = xl; x <= xh; x++) {
                    const int q[3] = {x + off[0], y + off[1], z + off[2]};
                    while (lo < P && less(xyz + 3 * (size_t)lo, q)) lo++;
                    if (lo < P && !less(q, xyz + 3 * (size_t)lo)) {
                        if (idx && tot < cap) idx[tot] = lo;
                        tot++;
                    }
                }
            }
        if (tot > (size_t)INT_MAX) return host_fail(-1, "too many mask entries");
        offsets[k + 1] = (int)tot;
    }
    if (idx && tot > cap) return host_fail(-1, "masks need %zu entries, capacity %zu", tot, cap);
    return 0;
}

extern "C" int rfba3b_spot_intensities_device(const int *d_xyz, int P, const double *d_phi, int K, double zfactor,
                                              double *d_spot, void *stream) {
    if (P < 0 || K < 0) return host_fail(-1, "P and K must be >= 0");
    if ((size_t)P * K == 0) return 0;
    if (!d_xyz || !d_phi || !d_spot) return host_fail(-1, "NULL buffer");
    if (!(zfactor > 0) || !std::isfinite(zfactor)) return host_fail(-1, "zfactor must be positive");
    if (int rc = host_dev_init(nullptr)) return rc;
    const size_t n = (size_t)P * K;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(b3_spot_intensity, dim3(grid), dim3(256), 0, (hipStream_t)stream, d_xyz, P, d_phi, K,
                       zfactor, d_spot);
    GHIP(hipGetLastError());
    return 0;
}
