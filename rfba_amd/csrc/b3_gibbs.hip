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
constexpr int kMaskChunk = 1024;  // mask pixels staged in LDS at a time
constexpr int kMtN = 624;         // MT19937 state words (mersenne.cpp MERS_N)

struct GibbsDev {
    int P, T, K, pad;
    const double *spot;   // [K][P]
    const int *moff;      // [K + 1]
    const int *midx;      // mask pixel indices
    uint8_t *sample;      // [K][T]
    double *resid;        // [T][P]
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
// b_on = e^min(x,0), b_off = e^min(-x,0).  alpha_t is the normalised forward
// message (forward_algorithm_delta2's delta_t minus a per-frame constant);
// backward sampling (forward_algorithm.h:542-566) draws the last frame from
// alpha_{T-1} and frame t from alpha_t(i) A[i][s_{t+1}], which is
// sample_unscaled_log's distribution (:520-532).
__device__ void ffbs_lane(const GibbsParams &G, int T, const double *ex, double *alpha, const uint32_t *ring,
                          int head, int RC, uint8_t *sn) {
    const double *A = G.A;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int t = 0; t < T; t++) {
        double p0, p1, p2;
        if (t == 0) {
            p0 = G.pi[0]; p1 = G.pi[1]; p2 = G.pi[2];
        } else {
            p0 = a0 * A[0] + a1 * A[3] + a2 * A[6];
            p1 = a0 * A[1] + a1 * A[4] + a2 * A[7];
            p2 = a0 * A[2] + a1 * A[5] + a2 * A[8];
        }
        const double x = ex[3 * t], bon = ex[3 * t + 1], boff = ex[3 * t + 2];
        const double n0 = p0 * bon, n1 = p1 * boff, n2 = p2 * boff;
        const double c = n0 + n1 + n2;
        if (c > 1e-250) {
            const double ic = 1.0 / c;
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
}

__global__ __launch_bounds__(kGibbsThreads) void b3_gibbs(GibbsParams G) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const GibbsDev c = G.chains[G.order[blockIdx.x]];
    const int T = c.T, P = c.P, K = c.K, Tc = G.Tcap;
    const int RC = Tc + kMtN;  // draw ring capacity

    double *alpha = (double *)lds;          // [Tc][3]
    double *ex = alpha + 3 * Tc;            // [Tc][3]
    double *dsum = ex + 3 * Tc;             // [Tc]
    double *muL = dsum + Tc;                // [kMaskChunk]
    int *idxL = (int *)(muL + kMaskChunk);  // [kMaskChunk]
    uint32_t *mt = (uint32_t *)(idxL + kMaskChunk);  // [kMtN]
    uint32_t *ring = mt + kMtN;             // [RC]
    uint8_t *so = (uint8_t *)(ring + RC);   // [Tc] state before the update
    uint8_t *sn = so + Tc;                  // [Tc] state after the update

    for (int i = tid; i < kMtN; i += kGibbsThreads) mt[i] = c.mt[i];
    const int pos0 = (int)c.mt[kMtN];
    __syncthreads();
    // Every thread tracks the ring's head and count; they change uniformly.
    int head = 0, cnt = kMtN - pos0;
    for (int i = tid; i < cnt; i += kGibbsThreads) ring[i] = mt_temper(mt[pos0 + i]);

    for (int s = 0; s < G.samples; s++) {
        const double i2v = 0.5 / G.var[s];
        for (int it = 0; it < G.iters; it++) {
            for (int k = 0; k < K; k++) {
                const int m0 = c.moff[k], m = c.moff[k + 1] - m0;
                const double *spk = c.spot + (size_t)k * P;
                uint8_t *smp = c.sample + (size_t)k * T;
                for (int t = tid; t < T; t += kGibbsThreads) {
                    so[t] = smp[t];
                    dsum[t] = 0.0;
                }
                __syncthreads();
                // 1. d_t over the mask, the spot taken out where it was on
                for (int c0 = 0; c0 < m; c0 += kMaskChunk) {
                    const int nc = min(kMaskChunk, m - c0);
                    for (int i = tid; i < nc; i += kGibbsThreads) {
                        const int p = c.midx[m0 + c0 + i];
                        idxL[i] = p;
                        muL[i] = spk[p];
                    }
                    __syncthreads();
                    for (int t = wave; t < T; t += kGibbsThreads / 64) {
                        const bool on = so[t] == 0;
                        const double *rt = c.resid + (size_t)t * P;
                        double acc = 0.0;
#pragma unroll 4
                        for (int i = lane; i < nc; i += 64) {
                            const double mu = muL[i];
                            double r = rt[idxL[i]];
                            if (on) r += mu;
                            acc += mu * (mu - 2.0 * r);
                        }
                        acc = wave_sum(acc);
                        if (lane == 0) dsum[t] += acc;
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
                if (tid == 0) ffbs_lane(G, T, ex, alpha, ring, head, RC, sn);
                head += T;
                head = head >= RC ? head - RC : head;
                cnt -= T;
                __syncthreads();
                // 3. the new sample, and the residual where on/off changed
                for (int t = tid; t < T; t += kGibbsThreads) smp[t] = sn[t];
                const bool restage = m > kMaskChunk;
                for (int c0 = 0; c0 < m; c0 += kMaskChunk) {
                    const int nc = min(kMaskChunk, m - c0);
                    if (restage) {
                        for (int i = tid; i < nc; i += kGibbsThreads) {
                            const int p = c.midx[m0 + c0 + i];
                            idxL[i] = p;
                            muL[i] = spk[p];
                        }
                        __syncthreads();
                    }
                    for (int t = wave; t < T; t += kGibbsThreads / 64) {
                        const bool was = so[t] == 0, now = sn[t] == 0;
                        if (was == now) continue;
                        double *rt = c.resid + (size_t)t * P;
                        for (int i = lane; i < nc; i += 64) {
                            const int p = idxL[i];
                            rt[p] = was ? rt[p] + muL[i] : rt[p] - muL[i];
                        }
                    }
                    if (restage) __syncthreads();
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
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nt; e += stride)
        c.resid[e] = (double)c.data[e] - c.base;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < ns; e += stride) c.sample[e] = 2;
}

// Mask checks of one chain per workgroup: offsets nondecreasing from 0,
// indices in [0, P) and strictly increasing within each spot.  out[2c] =
// number of violations, out[2c + 1] = mask entries.
__global__ void b3_gibbs_check(const GibbsDev *chains, int *out) {
    const GibbsDev c = chains[blockIdx.x];
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    int nb = 0;
    if (threadIdx.x == 0 && c.K > 0 && c.moff[0] != 0) nb++;
    for (int k = 0; k < c.K; k++) {
        const int a = c.moff[k], b = c.moff[k + 1];
        if (b < a) {
            nb += threadIdx.x == 0;
            continue;
        }
        for (int i = a + threadIdx.x; i < b; i += blockDim.x) {
            const int p = c.midx[i];
            if (p < 0 || p >= c.P) nb++;
            else if (i > a && c.midx[i - 1] >= p) nb++;
        }
    }
    if (nb) atomicAdd(&bad, nb);
    __syncthreads();
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = bad;
        out[2 * blockIdx.x + 1] = c.K > 0 ? c.moff[c.K] : 0;
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
    return 8 * (7 * (size_t)Tc + kMaskChunk) + 4 * (size_t)kMaskChunk + 4 * ((size_t)kMtN + Tc + kMtN) +
           2 * (size_t)Tc;
}

}  // namespace

struct rfba3b_gibbs {
    rfba3b_model model;
    std::vector<rfba3b_gibbs_chain> chains;
    std::vector<unsigned long long> draws;
    GibbsDev *d_chains = nullptr;
    int *d_order = nullptr;
    uint32_t *d_mt = nullptr;
    double *d_var = nullptr;
    int var_cap = 0;
    int Tmax = 0;
    size_t lds = 0;
    ~rfba3b_gibbs() {
        if (d_chains) (void)hipFree(d_chains);
        if (d_order) (void)hipFree(d_order);
        if (d_mt) (void)hipFree(d_mt);
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
        if ((size_t)c.P * c.T > (size_t)LLONG_MAX / 8) return host_fail(-1, "chain %d: T x P too large", i);
        if (!c.d_data || !c.d_resid || !c.d_sample) return host_fail(-1, "chain %d: data, residual or sample is NULL", i);
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
        d.P = c.P; d.T = c.T; d.K = c.K; d.pad = 0;
        d.spot = c.d_spot; d.moff = c.d_mask_off; d.midx = c.d_mask;
        d.sample = c.d_sample; d.resid = c.d_resid; d.samples = c.d_samples;
        d.data = c.d_data; d.base = c.base;
        d.mt = g->d_mt + (size_t)i * (kMtN + 1);
    }
    GHIP(hipMemcpy(g->d_mt, mts.data(), mts.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    GHIP(hipMalloc(&g->d_chains, n * sizeof(GibbsDev)));
    GHIP(hipMemcpy(g->d_chains, dev.data(), n * sizeof(GibbsDev), hipMemcpyHostToDevice));

    // mask checks on the device, then the launch order by work (T x mask entries)
    int *d_chk = nullptr;
    GHIP(hipMalloc(&d_chk, 2 * (size_t)n * sizeof(int)));
    hipLaunchKernelGGL(b3_gibbs_check, dim3(n), dim3(256), 0, nullptr, (const GibbsDev *)g->d_chains, d_chk);
    std::vector<int> chk(2 * (size_t)n);
    hipError_t e1 = hipGetLastError();
    hipError_t e2 = hipMemcpy(chk.data(), d_chk, chk.size() * sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(d_chk);
    GHIP(e1);
    GHIP(e2);
    std::vector<long long> work(n);
    for (int i = 0; i < n; i++) {
        if (chk[2 * i]) return host_fail(-1, "chain %d: %d invalid mask entries (offsets must start at 0 and not "
                                             "decrease; indices in [0, P), increasing within a spot)", i, chk[2 * i]);
        work[i] = (long long)chains[i].T * std::max(chk[2 * i + 1], 1) + (long long)chains[i].K * chains[i].T;
    }
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
