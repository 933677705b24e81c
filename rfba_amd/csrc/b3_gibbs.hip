// b3_gibbs.hip -- Gibbs sampling of the blink states of RFBA's 3B model on
// gfx950: GibbsSampler2::next (RFBA Main/sampled_multispot.h:681-709) for
// many chains at once, plus the two spot tables the sampler consumes,
// get_spot_pixels (multispot5.cc:1001-1020) and compute_spot_intensity
// (sampled_multispot.h:183-191).
//
// One workgroup runs one chain.  A spot update is
//   1. one pass over the spot's mask in every frame (pixels across lanes,
//      frames across waves) giving d_t = sum (r - mu)^2 - r^2, where r is
//      the residual with the spot removed; the reference's masked on/off
//      log-probabilities (spot_with_background.hh:120-147) differ by
//      B_on - B_off = -d_t / (2 var), and everything common to the three
//      states of a frame cancels in both the forward recursion and the
//      backward sampling;
//   2. the forward recursion on one lane over normalised messages (FP64),
//      then backward sampling with the chain's MT19937 draws;
//   3. a second pass that touches only the frames whose on/off status
//      changed, adding or removing the spot in the residual.
// The residual (data - sample intensities) lives in HBM in FP64; the mask's
// pixel indices and spot values are staged through LDS in chunks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "b3_internal.h"

namespace rfba3b {

constexpr int kGibbsThreads = 256;
constexpr int kMaskChunk = 2048;  // mask pixels staged in LDS at a time
constexpr int kMtN = 624;         // MT19937 state words (mersenne.cpp MERS_N)

typedef __attribute__((address_space(1))) double gdouble;

struct GibbsDev {
    int P, T, K, S;       // S: residual slots per frame (>= P)
    const double *spot;   // [K][P]
    const int *moff;      // [K + 1]
    const int *mpix;      // mask entries: pixel index (sorted by slot within a spot)
    const int *mslot;     // mask entries: residual slot
    const int *slot;      // [P] residual slot of each pixel
    uint8_t *sample;      // [K][T]
    double *resid;        // [T][S]
    uint8_t *samples;     // [S][K][T] or null
    const float *data;    // [T][P]
    double base;
    uint32_t *mt;         // [kMtN + 1]: state words, then the position of the next draw
};

struct GibbsParams {
    const GibbsDev *chains;
    const int *order;     // launch order (largest work first)
    int samples, iters, Tcap, pad;
    const double *var;    // [samples]
    double A[9], pi[3];
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
// in four dependent ranges: [0,227) reads only old words, [227,454) and
// [454,623) read the new words 227 below them, and 623 reads new[0], new[396].
__device__ void mt_twist(uint32_t *mt, int tid) {
    uint32_t v = 0;
    if (tid < 227) v = mt_word(mt[tid], mt[tid + 1], mt[tid + 397]);
    __syncthreads();
    if (tid < 227) mt[tid] = v;
    __syncthreads();
    if (tid < 227) v = mt_word(mt[tid + 227], mt[tid + 228], mt[tid]);
    __syncthreads();
    if (tid < 227) mt[tid + 227] = v;
    __syncthreads();
    if (tid < 169) v = mt_word(mt[tid + 454], mt[tid + 455], mt[tid + 227]);
    __syncthreads();
    if (tid < 169) mt[tid + 454] = v;
    __syncthreads();
    if (tid == 0) mt[623] = mt_word(mt[623], mt[0], mt[396]);
    __syncthreads();
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
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

// Forward recursion over normalised messages and backward sampling, on one
// lane.  ex[3t..3t+2] = (x_t, b_on, b_off) with x_t = B_on - B_off,
// b_on = e^min(x,0), b_off = e^min(-x,0).  alpha_t is the forward message
// up to a positive per-frame factor (forward_algorithm_delta2's delta_t
// minus a per-frame constant); backward sampling (forward_algorithm.h:
// 542-566) draws the last frame from alpha_{T-1} and frame t from
// alpha_t(i) A[i][s_{t+1}], which is sample_unscaled_log's distribution
// (:520-532) -- a per-frame factor changes neither.  The frames whose on/off
// status changes are listed in chg as (t << 1) | was_on; returns their count.
__device__ int ffbs_lane(const double *Api, int T, const double *ex, double *alpha, const uint32_t *ring,
                         int head, int RC, const uint8_t *so, uint8_t *sn, int *chg) {
    const double *A = Api, *pi = Api + 9;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int t = 0; t < T; t++) {
        double p0, p1, p2;
        if (t == 0) {
            p0 = pi[0]; p1 = pi[1]; p2 = pi[2];
        } else {
            p0 = a0 * A[0] + a1 * A[3] + a2 * A[6];
            p1 = a0 * A[1] + a1 * A[4] + a2 * A[7];
            p2 = a0 * A[2] + a1 * A[5] + a2 * A[8];
        }
        const double x = ex[3 * t], bon = ex[3 * t + 1], boff = ex[3 * t + 2];
        const double n0 = p0 * bon, n1 = p1 * boff, n2 = p2 * boff;
        const double c = n0 + n1 + n2;
        if (c > 1e-250) {
            // any positive scale will do: the hardware reciprocal
            const double ic = __builtin_amdgcn_rcp(c);
            a0 = n0 * ic; a1 = n1 * ic; a2 = n2 * ic;
        } else {
            // the same step in log space (the emission ratio underflowed)
            const double ninf = -__builtin_inf();
            const double l0 = (p0 > 0 ? log(p0) : ninf) + fmin(x, 0.0);
            const double l1 = (p1 > 0 ? log(p1) : ninf) + fmin(-x, 0.0);
            const double l2 = (p2 > 0 ? log(p2) : ninf) + fmin(-x, 0.0);
            const double mx = fmax(l0, fmax(l1, l2));
            if (mx == ninf) {
                a0 = a1 = a2 = 0.0;
            } else {
                const double e0 = exp(l0 - mx), e1 = exp(l1 - mx), e2 = exp(l2 - mx);
                const double is = 1.0 / (e0 + e1 + e2);
                a0 = e0 * is; a1 = e1 * is; a2 = e2 * is;
            }
        }
        alpha[3 * t] = a0;
        alpha[3 * t + 1] = a1;
        alpha[3 * t + 2] = a2;
    }
    const double scale = 1.0 / (65536.0 * 65536.0);
    int h = head;
    double u = (double)ring[h] * scale;
    h = (h + 1 == RC) ? 0 : h + 1;
    int next = select3(alpha[3 * (T - 1)], alpha[3 * (T - 1) + 1], alpha[3 * (T - 1) + 2], u);
    sn[T - 1] = (uint8_t)next;
    for (int t = T - 2; t >= 0; t--) {
        u = (double)ring[h] * scale;
        h = (h + 1 == RC) ? 0 : h + 1;
        next = select3(alpha[3 * t] * A[next], alpha[3 * t + 1] * A[3 + next], alpha[3 * t + 2] * A[6 + next], u);
        sn[t] = (uint8_t)next;
    }
    int n = 0;
    for (int t = 0; t < T; t++) {
        const bool was = so[t] == 0, now = sn[t] == 0;
        if (was != now) chg[n++] = (t << 1) | (was ? 1 : 0);
    }
    return n;
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

__global__ __launch_bounds__(kGibbsThreads) void b3_gibbs(GibbsParams G) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr int kWaves = kGibbsThreads / 64;
    const GibbsDev c = G.chains[G.order[blockIdx.x]];
    const int T = c.T, P = c.S, K = c.K, Tc = G.Tcap;  // P: the residual's row length
    const int RC = Tc + kMtN;  // draw ring capacity

    double *alpha = (double *)lds;          // [Tc][3]
    double *ex = alpha + 3 * Tc;            // [Tc][3]
    double *dsum = ex + 3 * Tc;             // [Tc]
    double *muL = dsum + Tc;                // [kMaskChunk]
    int *idxL = (int *)(muL + kMaskChunk);  // [kMaskChunk]
    int *chg = idxL + kMaskChunk;           // [Tc] frames whose on/off status changed
    int *nchg = chg + Tc;                   // [4]
    uint32_t *mt = (uint32_t *)(nchg + 4);  // [kMtN]
    uint32_t *ring = mt + kMtN;             // [RC]
    uint8_t *so = (uint8_t *)(ring + RC);   // [Tc] state before the update
    uint8_t *sn = so + Tc;                  // [Tc] state after the update
    __shared__ double Api[12];              // A[9], pi[3]
    // the residual through global-address-space pointers: rows are uniform
    // per wave, pixel offsets per lane
    gdouble *R = (gdouble *)c.resid;

    if (tid < 9) Api[tid] = G.A[tid];
    else if (tid < 12) Api[tid] = G.pi[tid - 9];
    for (int i = tid; i < kMtN; i += kGibbsThreads) mt[i] = c.mt[i];
    const int pos0 = (int)c.mt[kMtN];
    __syncthreads();
    // Every thread tracks the ring's head and count; they change uniformly.
    int head = 0, cnt = kMtN - pos0;
    for (int i = tid; i < cnt; i += kGibbsThreads) ring[i] = mt_temper(mt[pos0 + i]);
    const int nblk = (T + kFB - 1) / kFB;

    for (int s = 0; s < G.samples; s++) {
        const double i2v = 0.5 / G.var[s];
        for (int it = 0; it < G.iters; it++) {
            for (int k = 0; k < K; k++) {
                const int m0 = c.moff[k], m = c.moff[k + 1] - m0;
                const double *spk = c.spot + (size_t)k * c.P;
                uint8_t *smp = c.sample + (size_t)k * T;
                for (int t = tid; t < T; t += kGibbsThreads) {
                    so[t] = smp[t];
                    dsum[t] = 0.0;
                }
                __syncthreads();
                // 1. d_t over the mask, the spot taken out where it was on;
                //    each wave takes kFB consecutive frames at a time
                for (int c0 = 0; c0 < m; c0 += kMaskChunk) {
                    const int nc = min(kMaskChunk, m - c0);
                    for (int i = tid; i < nc; i += kGibbsThreads) {
                        idxL[i] = c.mslot[m0 + c0 + i];
                        muL[i] = spk[c.mpix[m0 + c0 + i]];
                    }
                    __syncthreads();
                    for (int b = wave; b < nblk; b += kWaves) {
                        const int t0 = b * kFB;
                        const gdouble *rt[kFB];
                        bool on[kFB];
                        double acc[kFB];
#pragma unroll
                        for (int f = 0; f < kFB; f++) {
                            const int t = min(t0 + f, T - 1);  // frames past T repeat the last, discarded
                            rt[f] = R + (size_t)t * P;
                            on[f] = so[t] == 0;
                            acc[f] = 0.0;
                        }
#pragma unroll 2
                        for (int i = lane; i < nc; i += 64) {
                            const int p = idxL[i];
                            const double mu = muL[i];
                            double r[kFB];
#pragma unroll
                            for (int f = 0; f < kFB; f++) r[f] = rt[f][p];
#pragma unroll
                            for (int f = 0; f < kFB; f++) {
                                const double r0 = on[f] ? r[f] + mu : r[f];
                                acc[f] += mu * (mu - 2.0 * r0);
                            }
                        }
                        const double v = wave_sum8(acc, lane);
                        const int t = t0 + ((lane >> 3) & 7);
                        if ((lane & 7) == 0 && t < T) dsum[t] += v;
                    }
                    __syncthreads();
                }
                for (int t = tid; t < T; t += kGibbsThreads) {
                    const double x = -dsum[t] * i2v;
                    ex[3 * t] = x;
                    ex[3 * t + 1] = x >= 0.0 ? 1.0 : exp(x);
                    ex[3 * t + 2] = x >= 0.0 ? exp(-x) : 1.0;
                }
                // 2. T draws, then the forward recursion and backward sampling
                while (cnt < T) {
                    mt_twist(mt, tid);
                    for (int i = tid; i < kMtN; i += kGibbsThreads) {
                        int w = head + cnt + i;
                        w = w >= RC ? w - RC : w;
                        ring[w] = mt_temper(mt[i]);
                    }
                    cnt += kMtN;
                }
                __syncthreads();
                if (tid == 0) nchg[0] = ffbs_lane(Api, T, ex, alpha, ring, head, RC, so, sn, chg);
                head += T;
                head = head >= RC ? head - RC : head;
                cnt -= T;
                __syncthreads();
                // 3. the new sample, and the residual in the frames whose
                //    on/off status changed (kFB of them per wave at a time)
                for (int t = tid; t < T; t += kGibbsThreads) smp[t] = sn[t];
                const int nch = nchg[0];
                const bool restage = m > kMaskChunk;
                if (nch > 0) {
                    for (int c0 = 0; c0 < m; c0 += kMaskChunk) {
                        const int nc = min(kMaskChunk, m - c0);
                        if (restage) {
                            for (int i = tid; i < nc; i += kGibbsThreads) {
                                idxL[i] = c.mslot[m0 + c0 + i];
                                muL[i] = spk[c.mpix[m0 + c0 + i]];
                            }
                            __syncthreads();
                        }
                        for (int j0 = wave * kFB; j0 < nch; j0 += kWaves * kFB) {
                            gdouble *rt[kFB];
                            double sg[kFB];
#pragma unroll
                            for (int f = 0; f < kFB; f++) {
                                const int e = __builtin_amdgcn_readfirstlane(chg[min(j0 + f, nch - 1)]);
                                rt[f] = R + (size_t)(e >> 1) * P;
                                sg[f] = j0 + f < nch ? ((e & 1) ? 1.0 : -1.0) : 0.0;  // +mu: spot removed
                            }
#pragma unroll 1
                            for (int i = lane; i < nc; i += 64) {
                                const int p = idxL[i];
                                const double mu = muL[i];
                                double r[kFB];
#pragma unroll
                                for (int f = 0; f < kFB; f++)
                                    if (sg[f] != 0.0) r[f] = rt[f][p];
#pragma unroll
                                for (int f = 0; f < kFB; f++)
                                    if (sg[f] != 0.0) rt[f][p] = r[f] + sg[f] * mu;
                            }
                        }
                        if (restage) __syncthreads();
                    }
                }
                __syncthreads();
            }
        }
        if (c.samples) {
            uint8_t *dst = c.samples + (size_t)s * K * T;
            for (int i = tid; i < K * T; i += kGibbsThreads) dst[i] = c.sample[i];
            __syncthreads();
        }
    }
    // the generator's state and the position of its next draw
    for (int i = tid; i < kMtN; i += kGibbsThreads) c.mt[i] = mt[i];
    if (tid == 0) c.mt[kMtN] = (uint32_t)(kMtN - cnt);
}

// GibbsSampler2::reset (sampled_multispot.h:430-438): all spots in state 2,
// sample intensities = base, i.e. residual = data - base.
__global__ void b3_gibbs_reset(const GibbsDev *chains) {
    const GibbsDev c = chains[blockIdx.y];
    const size_t nt = (size_t)c.T * c.P, ns = (size_t)c.K * c.T;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nt; e += stride) {
        const size_t t = e / c.P, p = e - t * c.P;
        c.resid[t * c.S + c.slot[p]] = (double)c.data[e] - c.base;
    }
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < ns; e += stride) c.sample[e] = 2;
}

// sample intensities (+ base) in the pixel list's order: data - residual
__global__ void b3_gibbs_export(const GibbsDev *chains, int i, double *out) {
    const GibbsDev c = chains[i];
    const size_t nt = (size_t)c.T * c.P;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nt; e += stride) {
        const size_t t = e / c.P, p = e - t * c.P;
        out[e] = (double)c.data[e] - c.resid[t * c.S + c.slot[p]];
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

size_t gibbs_lds(int Tc) {
    return 8 * (7 * (size_t)Tc + kMaskChunk) + 4 * ((size_t)kMaskChunk + Tc + 4) +
           4 * ((size_t)kMtN + Tc + kMtN) + 2 * (size_t)Tc;
}

}  // namespace

// Residual layout of a chain: with coordinates, voxel (x, y, z) of the
// pixels' bounding box goes to brick ((z/2 * NBY + y/2) * NBX + x/4), slot
// (z&1) * 8 + (y&1) * 4 + (x&3) within the brick's 16 (one 128-byte line of
// FP64); a region whose bricks would hold more than twice its pixels keeps
// the list order.  Returns the slots per frame.
static int brick_layout(const int *xyz, int P, std::vector<int> &slot) {
    slot.resize(P);
    if (xyz && P > 0) {
        int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
        for (int p = 0; p < P; p++)
            for (int j = 0; j < 3; j++) {
                lo[j] = std::min(lo[j], xyz[3 * p + j]);
                hi[j] = std::max(hi[j], xyz[3 * p + j]);
            }
        const long long nbx = ((long long)hi[0] - lo[0] + 4) / 4, nby = ((long long)hi[1] - lo[1] + 2) / 2,
                        nbz = ((long long)hi[2] - lo[2] + 2) / 2;
        const long long S = nbx * nby * nbz * 16;
        if (S <= 2LL * P + 4096) {
            for (int p = 0; p < P; p++) {
                const long long x = xyz[3 * p] - lo[0], y = xyz[3 * p + 1] - lo[1], z = xyz[3 * p + 2] - lo[2];
                slot[p] = (int)((((z >> 1) * nby + (y >> 1)) * nbx + (x >> 2)) * 16 + (z & 1) * 8 + (y & 1) * 4 + (x & 3));
            }
            return (int)S;
        }
    }
    for (int p = 0; p < P; p++) slot[p] = p;
    return P;
}

struct rfba3b_gibbs {
    rfba3b_model model;
    std::vector<rfba3b_gibbs_chain> chains;
    std::vector<GibbsDev> dev;
    std::vector<unsigned long long> draws;
    GibbsDev *d_chains = nullptr;
    int *d_order = nullptr;
    uint32_t *d_mt = nullptr;
    int *d_ints = nullptr;       // offsets, mask pixels, mask slots, pixel slots of every chain
    double *d_resid = nullptr;   // residuals of every chain
    double *d_var = nullptr;
    int var_cap = 0;
    int Tmax = 0;
    size_t lds = 0;
    ~rfba3b_gibbs() {
        if (d_chains) (void)hipFree(d_chains);
        if (d_order) (void)hipFree(d_order);
        if (d_mt) (void)hipFree(d_mt);
        if (d_ints) (void)hipFree(d_ints);
        if (d_resid) (void)hipFree(d_resid);
        if (d_var) (void)hipFree(d_var);
    }
};

static int gibbs_create(const rfba3b_gibbs_chain *chains, int n, const rfba3b_model *model, rfba3b_gibbs **out) {
    if (!out) return host_fail(-1, "plan pointer is NULL");
    *out = nullptr;
    if (!chains || n <= 0) return host_fail(-1, "no chains");
    if (int rc = host_check_model(model)) return rc;
    int ncu = 0;
    if (int rc = host_dev_init(&ncu)) return rc;
    int Tmax = 1;
    for (int i = 0; i < n; i++) {
        const rfba3b_gibbs_chain &c = chains[i];
        if (c.P <= 0 || c.T <= 0 || c.K < 0) return host_fail(-1, "chain %d: P and T must be positive, K >= 0", i);
        if (!c.d_data || !c.d_sample) return host_fail(-1, "chain %d: data or sample is NULL", i);
        if (c.K > 0 && (!c.d_spot || !c.d_mask_off || !c.d_mask))
            return host_fail(-1, "chain %d: spot table or masks are NULL", i);
        if (!std::isfinite(c.base)) return host_fail(-1, "chain %d: base is not finite", i);
        Tmax = std::max(Tmax, c.T);
    }
    const size_t lds = gibbs_lds(Tmax);
    if (lds > 160 * 1024) return host_fail(-1, "T = %d needs %zu bytes of LDS per chain (at most 160 KiB)", Tmax, lds);

    auto *g = new rfba3b_gibbs();
    std::unique_ptr<rfba3b_gibbs> own(g);
    g->model = *model;
    g->chains.assign(chains, chains + n);
    g->draws.resize(n);
    g->Tmax = Tmax;
    g->lds = lds;

    // masks to the host: checked, mapped to residual slots, sorted by slot
    std::vector<int> ints;                 // all chains' integer tables, concatenated
    std::vector<size_t> o_off(n), o_pix(n), o_slotm(n), o_slot(n), o_res(n);
    std::vector<int> S(n);
    std::vector<long long> work(n);
    size_t nres = 0;
    for (int i = 0; i < n; i++) {
        const rfba3b_gibbs_chain &c = chains[i];
        std::vector<int> off(c.K + 1, 0), mask, slot;
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
        S[i] = brick_layout(c.xyz, c.P, slot);
        if ((size_t)S[i] * c.T > (size_t)LLONG_MAX / 8) return host_fail(-1, "chain %d: T x P too large", i);
        o_off[i] = ints.size();
        ints.insert(ints.end(), off.begin(), off.end());
        o_pix[i] = ints.size();
        o_slotm[i] = o_pix[i] + mask.size();
        ints.resize(ints.size() + 2 * mask.size());
        for (int k = 0; k < c.K; k++) {
            std::vector<std::pair<int, int>> e;  // (slot, pixel)
            for (int j = off[k]; j < off[k + 1]; j++) e.emplace_back(slot[mask[j]], mask[j]);
            std::sort(e.begin(), e.end());
            for (size_t j = 0; j < e.size(); j++) {
                ints[o_pix[i] + off[k] + j] = e[j].second;
                ints[o_slotm[i] + off[k] + j] = e[j].first;
            }
        }
        o_slot[i] = ints.size();
        ints.insert(ints.end(), slot.begin(), slot.end());
        o_res[i] = nres;
        nres += ((size_t)S[i] * c.T + 31) & ~(size_t)31;  // 256-byte aligned chains
        work[i] = (long long)c.T * std::max(off[c.K], 1) + (long long)c.K * c.T;
    }
    GHIP(hipMalloc(&g->d_ints, std::max<size_t>(ints.size(), 1) * sizeof(int)));
    GHIP(hipMemcpy(g->d_ints, ints.data(), ints.size() * sizeof(int), hipMemcpyHostToDevice));
    GHIP(hipMalloc(&g->d_resid, nres * sizeof(double)));
    GHIP(hipMemset(g->d_resid, 0, nres * sizeof(double)));  // brick slots outside the region stay 0

    std::vector<uint32_t> mts((size_t)n * (kMtN + 1));
    std::vector<GibbsDev> &dev = g->dev;
    dev.resize(n);
    GHIP(hipMalloc(&g->d_mt, mts.size() * sizeof(uint32_t)));
    for (int i = 0; i < n; i++) {
        const rfba3b_gibbs_chain &c = chains[i];
        HostMt h(c.seed);
        h.discard(37ull + c.skip);
        memcpy(&mts[(size_t)i * (kMtN + 1)], h.mt, sizeof h.mt);
        mts[(size_t)i * (kMtN + 1) + kMtN] = (uint32_t)h.pos;
        g->draws[i] = c.skip;
        GibbsDev &d = dev[i];
        d.P = c.P; d.T = c.T; d.K = c.K; d.S = S[i];
        d.spot = c.d_spot;
        d.moff = g->d_ints + o_off[i];
        d.mpix = g->d_ints + o_pix[i];
        d.mslot = g->d_ints + o_slotm[i];
        d.slot = g->d_ints + o_slot[i];
        d.sample = c.d_sample; d.resid = g->d_resid + o_res[i]; d.samples = c.d_samples;
        d.data = c.d_data; d.base = c.base;
        d.mt = g->d_mt + (size_t)i * (kMtN + 1);
    }
    GHIP(hipMemcpy(g->d_mt, mts.data(), mts.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    GHIP(hipMalloc(&g->d_chains, n * sizeof(GibbsDev)));
    GHIP(hipMemcpy(g->d_chains, dev.data(), n * sizeof(GibbsDev), hipMemcpyHostToDevice));

    // launch order: largest work (T x mask entries) first
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
    GHIP(hipMalloc(&g->d_order, n * sizeof(int)));
    GHIP(hipMemcpy(g->d_order, order.data(), n * sizeof(int), hipMemcpyHostToDevice));

    if (lds > 64 * 1024)
        GHIP(hipFuncSetAttribute((const void *)b3_gibbs, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    GHIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)b3_gibbs, kGibbsThreads, lds));
    if (per_cu <= 0) return host_fail(-1, "Gibbs kernel does not fit on a CU (LDS %zu bytes)", lds);
    hipLaunchKernelGGL(b3_gibbs_reset, dim3(64, n), dim3(256), 0, nullptr, (const GibbsDev *)g->d_chains);
    GHIP(hipGetLastError());
    GHIP(hipDeviceSynchronize());
    *out = own.release();
    return 0;
}

extern "C" int rfba3b_gibbs_create(const rfba3b_gibbs_chain *chains, int n, const rfba3b_model *model,
                                   rfba3b_gibbs **plan) {
    std::lock_guard<std::mutex> lk(host_mutex());
    return gibbs_create(chains, n, model, plan);
}

extern "C" int rfba3b_gibbs_reset(rfba3b_gibbs *g, void *stream) {
    std::lock_guard<std::mutex> lk(host_mutex());
    if (!g) return host_fail(-1, "plan is NULL");
    const int n = (int)g->chains.size();
    hipLaunchKernelGGL(b3_gibbs_reset, dim3(64, n), dim3(256), 0, (hipStream_t)stream,
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
    hipLaunchKernelGGL(b3_gibbs, dim3((unsigned)g->chains.size()), dim3(kGibbsThreads), g->lds, st, P);
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
    hipLaunchKernelGGL(b3_gibbs_export, dim3(256), dim3(256), 0, (hipStream_t)stream, (const GibbsDev *)g->d_chains,
                       chain, d_out);
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
        for (int z = -r; z <= r; z++)
            for (int y = -r; y <= r; y++)
                for (int x = -r; x <= r; x++) {
                    const unsigned int m2 = unsigned(x * x) + unsigned(y * y) + unsigned(z * z);
                    if (!((float)m2 <= r2)) continue;
                    const int q[3] = {x + off[0], y + off[1], z + off[2]};
                    int lo = 0, hi = P;  // lower_bound
                    while (lo < hi) {
                        const int mid = lo + (hi - lo) / 2;
                        if (less(xyz + 3 * (size_t)mid, q)) lo = mid + 1;
                        else hi = mid;
                    }
                    if (lo < P && !less(q, xyz + 3 * (size_t)lo)) {
                        if (idx && tot < cap) idx[tot] = lo;
                        tot++;
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
