// b3_kernel.hip -- 3B spot log-likelihood and its adjoint on gfx950.
//
// Kernel A (b3_resid): for one (block, parameter set q, state sample s) and
// one frame t, the residual plane (FP64 sum, stored FP32)
//     R[t][p] = data[t][p] - (b0 + sum_j [state_s,j(t) == on] mu_j(p))
// over all K spots (sampled_multispot.h:83-97 add/remove, spot_shape
// storm.h:229-235).  mu is evaluated from separable per-axis tables.
//
// Kernel B (b3_eval): persistent work queue over items (block, q, s, k).
// For spot k, rb = R + [on_k(t)] mu_k is the data minus the other spots'
// sampled intensities and b0; per frame the workgroup forms
//     Soff = sum rb^2,  D = sum mu (2 rb - mu),  diff_n = sum dmu_n (rb - mu)
// (FP32 pixel terms, FP64 sums) which give the reference's per-frame log
// probabilities (spot_with_background.hh:127-147):
//     B(off) = B(bleached) = -Soff/(2 var) - P ln(2 pi var)/2,
//     B(on) = B(off) + D/(2 var),   d B(on)/d phi = diff / var.
// One lane then runs the HMM in FP64: a scaled forward recursion
// (forward_algorithm.h:192-237 in linear space, ln 0 -> 0 probability)
// keeping the normalised messages only every `ckpt` frames, and the
// adjoint: a backward recursion that recomputes each segment's forward
// messages from its checkpoint and accumulates
//     d logL / d phi = sum_t gamma_t(on) diff_t / var,
// the posterior-weighted form of the reference's forward-mode derivative
// (forward_algorithm.h:208-244).
#include <hip/hip_runtime.h>

#include "b3_internal.h"

namespace rfba3b {

__global__ void b3_reset(int *counter) {
    if (threadIdx.x == 0) *counter = 0;
}

// ---------------------------------------------------------------------------
// kernel A: residual planes
__global__ __launch_bounds__(256) void b3_resid(const DevBlock *__restrict__ blocks,
                                                const PlaneTask *__restrict__ tasks, int ntasks,
                                                float *__restrict__ R, const DevModel *__restrict__ mdl) {
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    for (int task = blockIdx.x; task < ntasks; task += gridDim.x) {
        const PlaneTask tk = tasks[task];
        const DevBlock &b = blocks[tk.block];
        const int X = b.X, Y = b.Y, Z = b.Z, K = b.K, T = b.T;
        const int P = X * Y * Z;
        const int n_ax = X + Y + Z;
        // tables: g[K][n_ax] (x | y | z entries), amp[K], on[K]
        float *g = lds;
        float *amp = g + K * n_ax;
        int *on = (int *)(amp + K);
        __syncthreads();
        const double zF = mdl->zfactor;
        for (int e = tid; e < K * n_ax; e += 256) {
            const int j = e / n_ax, a = e % n_ax;
            const double *ph = b.params + ((size_t)tk.q * K + j) * 5;
            const float sig = (float)ph[1];
            float d;
            if (a < X) d = (float)(a + b.ox) - (float)ph[2];
            else if (a < X + Y) d = (float)(a - X + b.oy) - (float)ph[3];
            else d = ((float)(a - X - Y + b.oz) - (float)ph[4]) / (float)zF;
            g[e] = expf(-(d * d) / (2.0f * sig * sig));
        }
        for (int j = tid; j < K; j += 256) {
            const double *ph = b.params + ((size_t)tk.q * K + j) * 5;
            amp[j] = (float)(ph[0] / (ph[1] * 2.5066282746310002));
            on[j] = b.states[((size_t)tk.s * K + j) * T + tk.t] == 0;
        }
        __syncthreads();
        const float *dt = b.data + (size_t)tk.t * P;
        float *Rt = R + tk.rslot + (size_t)tk.t * P;
        // FP32 spot images, FP64 accumulation, one rounding of the residual
        // (summing b0 + images in FP32 would drop the far tails of every
        // spot below half an ulp of b0: a one-sided bias).
        for (int p = tid; p < P; p += 256) {
            const int z = p / (X * Y), r = p - z * X * Y;
            const int y = r / X, x = r - y * X;
            double v = (double)dt[p] - b.b0;
            for (int j = 0; j < K; j++)
                if (on[j]) v -= (double)(amp[j] * g[j * n_ax + x] * g[j * n_ax + X + y] * g[j * n_ax + X + Y + z]);
            Rt[p] = (float)v;
        }
    }
}

// ---------------------------------------------------------------------------
// HMM in FP64 on one lane.

struct Hmm {
    double A[9], pi[3];
};

// emission ratios of frame t relative to the larger one: returns m and sets
// b_on, b_off in (0, 1], one of them exactly 1.
__device__ __forceinline__ void emis(double dlt, double &bon, double &boff, double &m) {
    if (dlt >= 0.0) {
        m = dlt;
        bon = 1.0;
        boff = exp(-dlt);
    } else {
        m = 0.0;
        bon = exp(dlt);
        boff = 1.0;
    }
}

// One normalised forward step: a <- normalise(b .* (a A)); returns log c.
__device__ __forceinline__ double fstep(const Hmm &h, double a[3], double bon, double boff, bool first) {
    double p0, p1, p2;
    if (first) {
        p0 = h.pi[0]; p1 = h.pi[1]; p2 = h.pi[2];
    } else {
        p0 = a[0] * h.A[0] + a[1] * h.A[3] + a[2] * h.A[6];
        p1 = a[0] * h.A[1] + a[1] * h.A[4] + a[2] * h.A[7];
        p2 = a[0] * h.A[2] + a[1] * h.A[5] + a[2] * h.A[8];
    }
    double n0 = p0 * bon, n1 = p1 * boff, n2 = p2 * boff;
    double c = n0 + n1 + n2;
    if (!(c > 1e-250)) {
        // underflow guard: the same step in log space
        const double l0 = log(p0) + log(bon), l1 = log(p1) + log(boff), l2 = log(p2) + log(boff);
        double mx = fmax(l0, fmax(l1, l2));
        if (mx == -__builtin_inf()) {
            a[0] = a[1] = a[2] = 0.0;
            return -__builtin_inf();
        }
        const double e0 = exp(l0 - mx), e1 = exp(l1 - mx), e2 = exp(l2 - mx);
        const double s = e0 + e1 + e2;
        a[0] = e0 / s; a[1] = e1 / s; a[2] = e2 / s;
        return mx + log(s);
    }
    const double ic = 1.0 / c;
    a[0] = n0 * ic; a[1] = n1 * ic; a[2] = n2 * ic;
    return log(c);
}

// ---------------------------------------------------------------------------
// kernel B: per-item emissions, logL, adjoint gradient

template <int MODE>  // MODE_LOGL, MODE_GRAD, MODE_DELTA
__global__ __launch_bounds__(256) void b3_eval(EvalParams E) {
    extern __shared__ double ldsd[];
    __shared__ int s_item;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const DevModel &md = *E.model;

    for (;;) {
        __syncthreads();
        if (tid == 0) s_item = atomicAdd(E.counter, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= E.nitems) break;
        const Item item = E.items[it];
        const DevBlock &b = E.blocks[item.block];
        const int X = b.X, Y = b.Y, Z = b.Z, T = b.T, K = b.K;
        const int XY = X * Y, P = XY * Z;
        const double *ph = b.params + ((size_t)item.q * K + item.k) * 5;
        const double I = ph[0], sg = ph[1], fx = ph[2], fy = ph[3], fz = ph[4];

        // LDS: em[T][8] doubles | xy table [XY][4] floats | z table [Z][4] floats | ckpt
        double *em = ldsd;
        float *xyt = (float *)(em + 8 * (size_t)E.Tmax);
        float *zt = xyt + 4 * E.XYmax;
        double *ck = (double *)(zt + 4 * E.Zmax + (E.Zmax & 1) * 4);

        const float Is = (float)I, ss = (float)sg;
        const float inv = (float)(1.0 / (sg * 2.5066282746310002));
        const float i2s2 = (float)(1.0 / (2.0 * sg * sg));
        const float Iis2 = (float)(I / (sg * sg));
        const float zF = (float)md.zfactor;
        const float zdiv = md.z_exact ? zF * zF : 1.0f;
        for (int e = tid; e < XY; e += 256) {
            const int y = e / X, x = e - y * X;
            const float dx = (float)(x + b.ox) - (float)fx, dy = (float)(y + b.oy) - (float)fy;
            const float sx = -(dx * dx) * i2s2, sy = -(dy * dy) * i2s2;
            float *q = xyt + 4 * e;
            q[0] = expf(sx) * expf(sy) * inv;  // e^(sx+sy) / (sigma sqrt(2 pi))
            q[1] = sx + sy;
            q[2] = dx * Iis2;
            q[3] = dy * Iis2;
        }
        for (int z = tid; z < Z; z += 256) {
            const float dz = (float)(z + b.oz) - (float)fz;
            const float dzs = dz / zF;
            const float sz = -(dzs * dzs) * i2s2;
            float *q = zt + 4 * z;
            q[0] = expf(sz);
            q[1] = sz;
            q[2] = dz * Iis2 / zdiv;
            q[3] = 0.0f;
        }
        __syncthreads();

        // ---- per-frame sums: wave w takes frames w, w+4, ...
        const float *Rb = E.R + item.rslot;
        const uint8_t *mask = b.mask;
        const uint8_t *stk = b.states + ((size_t)item.s * K + item.k) * T;
        const float msig = -Is / ss;
        for (int t = wave; t < T; t += 4) {
            const bool onk = (E.rmode == 0) && stk[t] == 0;
            const float *Rt = Rb + (size_t)t * P;
            double aS = 0, aD = 0, a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
            for (int c0 = 0; c0 < XY; c0 += 64) {
                const int xy = c0 + lane;
                if (xy < XY) {
                    const float4 q = *(const float4 *)(xyt + 4 * xy);
                    float fS = 0, fD = 0, f0 = 0, f1 = 0, f2 = 0, f3 = 0, f4 = 0;
                    for (int z = 0; z < Z; z++) {
                        if (mask && !mask[(size_t)z * XY + xy]) continue;
                        const float4 w = *(const float4 *)(zt + 4 * z);
                        const float Ev = q.x * w.x;
                        const float mu = Is * Ev;
                        const float s = q.y + w.y;
                        const float r = Rt[(size_t)z * XY + xy];
                        // rb = data - other spots - b0, e = rb - mu, each
                        // rounded once from the residual r
                        const float rb = onk ? r + mu : r;
                        const float e = onk ? r : r - mu;
                        fS += rb * rb;
                        fD += mu * (rb + e);
                        if (MODE == MODE_GRAD) {
                            f0 += Ev * e;
                            f1 += Ev * msig * (1.0f + 2.0f * s) * e;
                            f2 += Ev * q.z * e;
                            f3 += Ev * q.w * e;
                            f4 += Ev * w.z * e;
                        }
                    }
                    aS += fS; aD += fD;
                    if (MODE == MODE_GRAD) { a0 += f0; a1 += f1; a2 += f2; a3 += f3; a4 += f4; }
                }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                aS += __shfl_xor(aS, o);
                aD += __shfl_xor(aD, o);
                if (MODE == MODE_GRAD) {
                    a0 += __shfl_xor(a0, o);
                    a1 += __shfl_xor(a1, o);
                    a2 += __shfl_xor(a2, o);
                    a3 += __shfl_xor(a3, o);
                    a4 += __shfl_xor(a4, o);
                }
            }
            if (lane == 0) {
                double *m = em + 8 * t;
                m[0] = aS; m[1] = aD; m[2] = a0; m[3] = a1; m[4] = a2; m[5] = a3; m[6] = a4;
            }
        }
        __syncthreads();

        // ---- HMM on one lane
        if (tid == 0) {
            Hmm h;
            for (int i = 0; i < 9; i++) h.A[i] = md.A[i];
            for (int i = 0; i < 3; i++) h.pi[i] = md.pi[i];
            const double i2v = 0.5 / md.variance;
            const double cst = (double)b.npix * log(2.0 * M_PI * md.variance) / 2.0;
            const int ckpt = E.ckpt > 0 ? min(E.ckpt, T) : (int)ceil(sqrt((double)T));
            double a[3] = {0, 0, 0};
            double L = 0.0;
            double *dout = (MODE == MODE_DELTA) ? E.delta_out + (size_t)it * 3 * T : nullptr;
            for (int t = 0; t < T; t++) {
                const double e1 = -em[8 * t] * i2v - cst;
                double bon, boff, m;
                emis(em[8 * t + 1] * i2v, bon, boff, m);
                if (t % ckpt == 0) {  // message entering frame t
                    ck[3 * (t / ckpt)] = a[0];
                    ck[3 * (t / ckpt) + 1] = a[1];
                    ck[3 * (t / ckpt) + 2] = a[2];
                }
                if (MODE == MODE_DELTA) {
                    // delta_t(j) = ln alpha_t(j) = ln pred_j + B(j, t) + logL_{t-1}
                    // (forward_algorithm.h:344-382), evaluated in log space so
                    // that states whose scaled weight underflows keep their
                    // value; ln 0 = -1e100 as in the reference (:34-40).
                    double pr[3];
                    if (t == 0) {
                        pr[0] = h.pi[0]; pr[1] = h.pi[1]; pr[2] = h.pi[2];
                    } else {
                        pr[0] = a[0] * h.A[0] + a[1] * h.A[3] + a[2] * h.A[6];
                        pr[1] = a[0] * h.A[1] + a[1] * h.A[4] + a[2] * h.A[7];
                        pr[2] = a[0] * h.A[2] + a[1] * h.A[5] + a[2] * h.A[8];
                    }
                    const double Bon = e1 + em[8 * t + 1] * i2v;
                    dout[3 * t] = (pr[0] > 0 ? log(pr[0]) : -1e100) + Bon + L;
                    dout[3 * t + 1] = (pr[1] > 0 ? log(pr[1]) : -1e100) + e1 + L;
                    dout[3 * t + 2] = (pr[2] > 0 ? log(pr[2]) : -1e100) + e1 + L;
                }
                L += fstep(h, a, bon, boff, t == 0) + e1 + m;
            }
            const size_t oi = ((size_t)item.q * K + item.k) * b.S + item.s;
            b.logL[oi] = L;
            if (MODE == MODE_GRAD) {
                // adjoint: beta_t = A (b_{t+1} .* beta_{t+1}) / c_{t+1}; segments
                // from the last, each re-running the forward steps from its
                // checkpoint to recover alpha_t and c_t.
                double g[5] = {0, 0, 0, 0, 0};
                double beta[3] = {1.0, 1.0, 1.0};
                double *sa = ck + 3 * ((T + ckpt - 1) / ckpt);  // segment alphas [ckpt][3]
                double *sc = sa + 3 * ckpt;                    // segment log c   [ckpt]
                double c_next = 0.0;                           // log c_{t1}
                double bon_n = 0, boff_n = 0;                  // b(t1)
                const int nseg = (T + ckpt - 1) / ckpt;
                for (int sgi = nseg - 1; sgi >= 0; sgi--) {
                    const int t0 = sgi * ckpt, t1 = min(T, t0 + ckpt);
                    double ar[3] = {ck[3 * sgi], ck[3 * sgi + 1], ck[3 * sgi + 2]};
                    for (int t = t0; t < t1; t++) {
                        double bon, boff, m;
                        emis(em[8 * t + 1] * i2v, bon, boff, m);
                        sc[t - t0] = fstep(h, ar, bon, boff, t == 0);
                        sa[3 * (t - t0)] = ar[0];
                        sa[3 * (t - t0) + 1] = ar[1];
                        sa[3 * (t - t0) + 2] = ar[2];
                    }
                    for (int t = t1 - 1; t >= t0; t--) {
                        if (t < T - 1) {
                            // beta_t(i) = sum_j A_ij b_j(t+1) beta_{t+1}(j) / c_{t+1}
                            const double ic = exp(-c_next);
                            const double w0 = bon_n * beta[0], w1 = boff_n * beta[1], w2 = boff_n * beta[2];
                            const double n0 = (h.A[0] * w0 + h.A[1] * w1 + h.A[2] * w2) * ic;
                            const double n1 = (h.A[3] * w0 + h.A[4] * w1 + h.A[5] * w2) * ic;
                            const double n2 = (h.A[6] * w0 + h.A[7] * w1 + h.A[8] * w2) * ic;
                            beta[0] = n0; beta[1] = n1; beta[2] = n2;
                        }
                        const double gam = sa[3 * (t - t0)] * beta[0];
                        const double *mm = em + 8 * t;
                        for (int n = 0; n < 5; n++) g[n] += gam * mm[2 + n];
                        double m;
                        emis(mm[1] * i2v, bon_n, boff_n, m);
                        c_next = sc[t - t0];  // log c_t (emission shift m excluded)
                    }
                }
                const double iv = 1.0 / md.variance;
                for (int n = 0; n < 5; n++) b.grad[oi * 5 + n] = g[n] * iv;
            }
        }
    }
}

template __global__ void b3_eval<MODE_LOGL>(EvalParams);
template __global__ void b3_eval<MODE_GRAD>(EvalParams);
template __global__ void b3_eval<MODE_DELTA>(EvalParams);

}  // namespace rfba3b
