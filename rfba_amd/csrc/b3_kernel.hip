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
// One lane then runs the HMM in FP64: a normalised forward recursion in
// extended range, equivalent to the log-space one (forward_algorithm.h:
// 192-237; ln 0 -> probability 0)
// keeping the normalised messages only every `ckpt` frames, and the
// adjoint: a backward recursion that recomputes each segment's forward
// messages from its checkpoint and accumulates
//     d logL / d phi = sum_t gamma_t(on) diff_t / var,
// the posterior-weighted form of the reference's forward-mode derivative
// (forward_algorithm.h:208-244).
#include <hip/hip_runtime.h>

#include "b3_internal.h"
#include "b3_wave.h"

namespace rfba3b {

__global__ void b3_reset(int *counter) {
    if (threadIdx.x == 0) *counter = 0;
}

// ---------------------------------------------------------------------------
// kernel A: residual planes
__global__ __launch_bounds__(256) void b3_resid(const DevBlock *__restrict__ blocks,
                                                const PlaneTask *__restrict__ tasks, int ntasks,
                                                float *__restrict__ R, double *__restrict__ S0, int s0_stride,
                                                const DevModel *__restrict__ mdl) {
    extern __shared__ float lds[];
    __shared__ double red[4];
    const int tid = threadIdx.x;
    for (int task = blockIdx.x; task < ntasks; task += gridDim.x) {
        const PlaneTask tk = tasks[task];
        const DevBlock &b = blocks[tk.block];
        const int X = b.X, Y = b.Y, Z = b.Z, K = b.K, T = b.T;
        const int P = X * Y * Z;
        // tables: gxy[K][X+Y] (amp * x factor | y factor, padded to 4 floats),
        // gz[K][Zp] (z factors, Zp = Z rounded up to 8, zero past Z: 16-byte
        // aligned rows read as two float4 per 8-deep run), amp[K], and the
        // on-states of this frame as 64-bit masks (two words each)
        const int nxy = X + Y, nxyp = (K * nxy + 3) & ~3, Zp = (Z + 7) & ~7;
        float *gxy = lds;
        float *gzt = gxy + nxyp;
        float *amp = gzt + K * Zp;
        unsigned *onm = (unsigned *)(amp + K);  // [ceil(K / 64)][2]
        __syncthreads();
        const double zF = mdl->zfactor;
        for (int j = tid; j < K; j += 256) {
            const double *ph = b.params + ((size_t)tk.q * K + j) * 5;
            amp[j] = (float)(ph[0] / (ph[1] * 2.5066282746310002));
        }
        // on-state masks of this frame: wave 0, 64 spots per ballot
        if (tid < 64) {
            for (int j0 = 0; j0 < K; j0 += 64) {
                const int j = j0 + tid;
                const bool on = j < K && b.states[((size_t)tk.s * K + j) * T + tk.t] == 0;
                const unsigned long long m = __ballot(on);
                if (tid == 0) {
                    onm[2 * (j0 >> 6)] = (unsigned)m;
                    onm[2 * (j0 >> 6) + 1] = (unsigned)(m >> 32);
                }
            }
        }
        __syncthreads();
        for (int e = tid; e < K * nxy; e += 256) {
            const int j = e / nxy, a = e - j * nxy;
            const double *ph = b.params + ((size_t)tk.q * K + j) * 5;
            const float sig = (float)ph[1];
            const float d = a < X ? (float)(a + b.ox) - (float)ph[2] : (float)(a - X + b.oy) - (float)ph[3];
            const float g = expf(-(d * d) / (2.0f * sig * sig));
            gxy[e] = a < X ? amp[j] * g : g;
        }
        for (int e = tid; e < K * Zp; e += 256) {
            const int j = e / Zp, z = e - j * Zp;
            float v = 0.0f;
            if (z < Z) {
                const double *ph = b.params + ((size_t)tk.q * K + j) * 5;
                const float sig = (float)ph[1];
                const float d = ((float)(z + b.oz) - (float)ph[4]) / (float)zF;
                v = expf(-(d * d) / (2.0f * sig * sig));
            }
            gzt[e] = v;
        }
        __syncthreads();
        const float *dt = b.data + (size_t)tk.t * P;
        float *Rt = R + tk.rslot + (size_t)tk.t * P;
        // FP32 spot images summed in FP32 (relative to their own sum), then
        // data - b0 - sum in FP64 and one rounding of the residual (adding
        // b0 into the FP32 sum would drop the far tails of every spot below
        // half an ulp of b0: a one-sided bias).
        // threads over (x,y) column x 8-deep z runs: the spot's xy factor is
        // formed once per run; the product order ((amp gx) gy) gz is kept.
        double ss = 0.0;
        constexpr int RZ = 8;
        const int XY = X * Y, nzc = (Z + RZ - 1) / RZ;
        for (int e = tid; e < XY * nzc; e += 256) {
            const int zc = e / XY, col = e - zc * XY;
            const int y = col / X, x = col - y * X;
            const int z0 = zc * RZ;
            // data first: the loads are in flight during the spot sum
            float dv[RZ], m[RZ];
#pragma unroll
            for (int u = 0; u < RZ; u++) {
                dv[u] = dt[(size_t)min(z0 + u, Z - 1) * XY + col];
                m[u] = 0.0f;
            }
            // the on-spots in increasing order, from the masks in scalar
            // registers (the sum's order is that of the spots)
            for (int j0 = 0; j0 < K; j0 += 64) {
                unsigned long long msk =
                    (unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)onm[2 * (j0 >> 6)]) |
                    ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)onm[2 * (j0 >> 6) + 1]) << 32);
                while (msk) {
                    const int j = j0 + __builtin_ctzll(msk);
                    msk &= msk - 1ull;
                    const float *gj = gxy + j * nxy;
                    const float a = gj[x] * gj[X + y];  // (amp gx) gy
                    const float4 *gz = (const float4 *)(gzt + j * Zp + z0);
                    const float4 g0 = gz[0], g1 = gz[1];
                    m[0] += a * g0.x; m[1] += a * g0.y; m[2] += a * g0.z; m[3] += a * g0.w;
                    m[4] += a * g1.x; m[5] += a * g1.y; m[6] += a * g1.z; m[7] += a * g1.w;
                }
            }
#pragma unroll
            for (int u = 0; u < RZ; u++)
                if (z0 + u < Z) {
                    const double v = ((double)dv[u] - b.b0) - (double)m[u];
                    const float rf = (float)v;
                    Rt[(size_t)(z0 + u) * XY + col] = rf;
                    ss += (double)rf * (double)rf;
                }
        }
        // per-frame sum of squares of the stored residual (exact products)
        if (S0) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);
            if ((tid & 63) == 0) red[tid >> 6] = ss;
            __syncthreads();
            if (tid == 0) S0[(size_t)tk.plane * s0_stride + tk.t] = (red[0] + red[1]) + (red[2] + red[3]);
        }
    }
}

// ---------------------------------------------------------------------------
// HMM in FP64 on one lane, in extended range (a mantissa and an exponent
// per value, xsplit): the reference runs the recursion in log space
// (forward_algorithm.h:192-237), where a state whose weight is e^-10000
// relative to another keeps it; normalised doubles would lose it.

struct Hmm {
    double A[9], pi[3];
};

// Emission ratios of a frame relative to the larger one: b_on =
// e^min(dlt,0), b_off = e^min(-dlt,0); the larger log emission, max(dlt, 0),
// is added to logL separately.
struct Emis {
    double onm, one, offm, offe;
};

__device__ __forceinline__ Emis xemis(double dlt) {
    Emis b;
    xexp(fmin(dlt, 0.0), b.onm, b.one);
    xexp(fmin(-dlt, 0.0), b.offm, b.offe);
    return b;
}

// The predicted message p = a A (pi in frame 0).
__device__ __forceinline__ void xpred(const Hmm &h, const double am[3], const double ae[3], bool first, double pm[3],
                                      double pe[3]) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
        if (first)
            xsplit(h.pi[j], 0.0, pm[j], pe[j]);
        else
            xsum3(am[0] * h.A[j], ae[0], am[1] * h.A[3 + j], ae[1], am[2] * h.A[6 + j], ae[2], pm[j], pe[j]);
    }
}

// a <- (b .* p) / c with c = sum(b .* p), returned as (cm, ce); returns ln c
// (-inf when every state has probability 0).
__device__ __forceinline__ double xupdate(const double pm[3], const double pe[3], const Emis &b, double am[3],
                                          double ae[3], double &cm, double &ce) {
    double nm[3], ne[3];
    xsplit(pm[0] * b.onm, pe[0] + b.one, nm[0], ne[0]);
    xsplit(pm[1] * b.offm, pe[1] + b.offe, nm[1], ne[1]);
    xsplit(pm[2] * b.offm, pe[2] + b.offe, nm[2], ne[2]);
    xsum3(nm[0], ne[0], nm[1], ne[1], nm[2], ne[2], cm, ce);
    if (!(cm > 0.0)) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            am[j] = 0.0;
            ae[j] = -__builtin_inf();
        }
        return -__builtin_inf();
    }
    const double ic = 1.0 / cm;
#pragma unroll
    for (int j = 0; j < 3; j++) xsplit(nm[j] * ic, ne[j] - ce, am[j], ae[j]);
    return log(cm) + ce * M_LN2;
}

// One normalised forward step; returns ln c.
__device__ __forceinline__ double xfstep(const Hmm &h, double am[3], double ae[3], const Emis &b, bool first,
                                         double &cm, double &ce) {
    double pm[3], pe[3];
    xpred(h, am, ae, first, pm, pe);
    return xupdate(pm, pe, b, am, ae, cm, ce);
}

// beta_t(i) = sum_j A_ij b_j(t+1) beta_{t+1}(j) / c_{t+1}, from the next
// frame's emissions b and normaliser c.
__device__ __forceinline__ void xbeta(const Hmm &h, double bm[3], double be[3], const Emis &b, double cm, double ce) {
    const double wm0 = b.onm * bm[0], we0 = b.one + be[0];
    const double wm1 = b.offm * bm[1], we1 = b.offe + be[1];
    const double wm2 = b.offm * bm[2], we2 = b.offe + be[2];
    const double ic = cm > 0.0 ? 1.0 / cm : 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double nm, ne;
        xsum3(h.A[3 * i] * wm0, we0, h.A[3 * i + 1] * wm1, we1, h.A[3 * i + 2] * wm2, we2, nm, ne);
        xsplit(nm * ic, ne - ce, bm[i], be[i]);
    }
}

// m 2^e as a double (a posterior weight, at most 1 up to rounding)
__device__ __forceinline__ double xvalue(double m, double e) { return ldexp(m, (int)fmin(fmax(e, -2000.0), 64.0)); }

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
            const double ninf = -__builtin_inf();
            double am[3] = {0.0, 0.0, 0.0}, ae[3] = {ninf, ninf, ninf};
            double L = 0.0;
            double *dout = (MODE == MODE_DELTA) ? E.delta_out + (size_t)it * 3 * T : nullptr;
            for (int t = 0; t < T; t++) {
                const double e1 = -em[8 * t] * i2v - cst;
                const double dlt = em[8 * t + 1] * i2v;
                if (t % ckpt == 0) {  // message entering frame t
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        ck[6 * (t / ckpt) + j] = am[j];
                        ck[6 * (t / ckpt) + 3 + j] = ae[j];
                    }
                }
                double pm[3], pe[3];
                xpred(h, am, ae, t == 0, pm, pe);
                if (MODE == MODE_DELTA) {
                    // delta_t(j) = ln pred_j + B(j, t) + logL_{t-1}
                    // (forward_algorithm.h:344-382); ln 0 = -1e100 as in the
                    // reference (:34-40).
                    const double Bon = e1 + dlt;
                    dout[3 * t] = (pm[0] > 0 ? log(pm[0]) + pe[0] * M_LN2 : -1e100) + Bon + L;
                    dout[3 * t + 1] = (pm[1] > 0 ? log(pm[1]) + pe[1] * M_LN2 : -1e100) + e1 + L;
                    dout[3 * t + 2] = (pm[2] > 0 ? log(pm[2]) + pe[2] * M_LN2 : -1e100) + e1 + L;
                }
                double cm, ce;
                L += xupdate(pm, pe, xemis(dlt), am, ae, cm, ce) + e1 + fmax(dlt, 0.0);
            }
            const size_t oi = ((size_t)item.q * K + item.k) * b.S + item.s;
            b.logL[oi] = L;
            if (MODE == MODE_GRAD) {
                // adjoint: beta_t = A (b_{t+1} .* beta_{t+1}) / c_{t+1}; segments
                // from the last, each re-running the forward steps from its
                // checkpoint to recover alpha_t(on) and c_t.
                double g[5] = {0, 0, 0, 0, 0};
                double bm[3] = {0.5, 0.5, 0.5}, be[3] = {1.0, 1.0, 1.0};  // beta_{T-1} = 1
                const int nseg = (T + ckpt - 1) / ckpt;
                double *sa = ck + 6 * nseg;  // segment alpha_t(on) [ckpt][2] (m, e)
                double *sc = sa + 2 * ckpt;  // segment c_t        [ckpt][2]
                Emis bn = xemis(0.0);        // b(t+1)
                double cnm = 1.0, cne = 0.0; // c_{t+1}
                for (int sgi = nseg - 1; sgi >= 0; sgi--) {
                    const int t0 = sgi * ckpt, t1 = min(T, t0 + ckpt);
                    double arm[3], are[3];
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        arm[j] = ck[6 * sgi + j];
                        are[j] = ck[6 * sgi + 3 + j];
                    }
                    for (int t = t0; t < t1; t++) {
                        double cm, ce;
                        xfstep(h, arm, are, xemis(em[8 * t + 1] * i2v), t == 0, cm, ce);
                        sa[2 * (t - t0)] = arm[0];
                        sa[2 * (t - t0) + 1] = are[0];
                        sc[2 * (t - t0)] = cm;
                        sc[2 * (t - t0) + 1] = ce;
                    }
                    for (int t = t1 - 1; t >= t0; t--) {
                        if (t < T - 1) xbeta(h, bm, be, bn, cnm, cne);
                        const double gam = xvalue(sa[2 * (t - t0)] * bm[0], sa[2 * (t - t0) + 1] + be[0]);
                        const double *mm = em + 8 * t;
                        for (int n = 0; n < 5; n++) g[n] += gam * mm[2 + n];
                        bn = xemis(mm[1] * i2v);
                        cnm = sc[2 * (t - t0)];
                        cne = sc[2 * (t - t0) + 1];
                    }
                }
                const double iv = 1.0 / md.variance;
                for (int n = 0; n < 5; n++) b.grad[oi * 5 + n] = g[n] * iv;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Kernel C (b3_corr_mfma): the per-frame sums of kernel B for all spots of a
// residual plane at once.  With rb = R + [on_k(t)] mu_k,
//     Soff = S0 + [on] (2 C + Q),   D = 2 C + (on ? Q : -Q),
//     diff_n = G_n - [off] H_n,
// where S0(t) = sum R^2 (from kernel A), C(t) = sum mu R, G_n(t) = sum dmu_n R
// and the frame-independent Q = sum mu^2, H_n = sum dmu_n mu.  C and G are a
// product of the plane [T x P] with the spots' basis [P x 6 nk]: the
// workgroup streams 32-pixel chunks of the plane (128 frames per pass) and
// of the basis through LDS, so the plane is read once per 8 spots.
// The product runs on the matrix cores with v_mfma_f32_16x16x4_f32
// (lane l holds A[l%16][l/16] = R[frame][pixel] and B[l/16][l%16] =
// basis[pixel][column]; result rows 4(l/16)+r, column l%16).  Passes of 128
// frames; wave w owns frame tiles 2w, 2w+1 (16 frames each) x all column
// tiles; FP32 matrix accumulation over four 32-pixel chunks, FP64 beyond.
// Each lane writes the em fields of its own (frame, column) results.
typedef float v4f __attribute__((ext_vector_type(4)));

// FUNC: the resident functor's form (a region mask, pixel splits with raw
// partial sums); the batched form compiles without them.  NCU > 0: matrix
// work on the first NCU 16-column tiles only (tasks whose spots' basis
// columns all lie there: at most 2 spots for the gradient); the sums of
// those columns are the same bits as with all tiles.
template <bool GRAD, bool FUNC, int NCU>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void b3_corr_mfma(CorrParams C) {
    extern __shared__ double ldsc[];
    constexpr int NB = GRAD ? 6 : 1;
    constexpr int NCOL = kCorrSpots * NB;   // 48 or 8 basis columns
    constexpr int NCT = (NCOL + 15) / 16;   // 16-column tiles
    // basis chunk row stride: a multiple of 16 that is 16 or 48 mod 64, so the
    // four pixel rows a B-operand read touches fall on disjoint bank groups
    constexpr int BSTR = NCT * 16;
    constexpr int NCX = NCU > 0 && NCU < NCT ? NCU : NCT;  // column tiles with matrix work
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // Workgroup b runs on XCD b % 8: give each XCD a contiguous range of
    // tasks so that the spot groups of one plane share that XCD's L2 (the
    // grid is padded to a multiple of 8).
    const int per_xcd = gridDim.x >> 3;
    const int vtask = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    const int nsplit = FUNC && C.nsplit > 1 ? C.nsplit : 1;
    if (vtask >= C.ntasks * nsplit) return;
    const CorrTask tk = C.tasks[vtask / nsplit];
    const int split = vtask - (vtask / nsplit) * nsplit;
    const PlaneRef pr = C.planes[tk.plane];
    const DevBlock &b = C.blocks[pr.block];
    const int X = b.X, Y = b.Y, Z = b.Z, T = b.T, K = b.K;
    const int XY = X * Y, P = XY * Z, nax = X + Y + Z;
    const int nk = tk.nk;
    const DevModel &md = *C.model;
    // pixel range of this (split) task
    const int pbeg = FUNC && nsplit > 1 ? split * C.pstep : 0;
    const int pend = FUNC && nsplit > 1 ? min(P, pbeg + C.pstep) : P;

    double *qh = ldsc;                                   // [8][6] Q, H_0..4
    double *qw = qh + kCorrSpots * 6;                    // [4 waves][8][6] partial Q, H
    float *Rs = (float *)(qw + 4 * kCorrSpots * 6);      // [kCorrFrames][kCorrRow]
    float *Bs = Rs + kCorrFrames * kCorrRow;             // [kCorrPix][BSTR]
    float *ax = Bs + kCorrPix * BSTR;                    // [3][nax][8]

    // Axis tables, component c of axis entry a (x, then y, then z) for spot j
    // at ax[(c nax + a) 8 + j]: the 8 spots x 8 consecutive pixels a wave
    // reads are 64 consecutive words.
    const float zF = (float)md.zfactor;
    const float zdiv = md.z_exact ? zF * zF : 1.0f;
    for (int e = tid; e < nk * nax; e += 256) {
        const int a = e / nk, j = e - a * nk;
        const double *ph = b.params + ((size_t)pr.q * K + tk.k0 + j) * 5;
        const double sg = ph[1];
        const float i2s2 = (float)(1.0 / (2.0 * sg * sg));
        const float Iis2 = (float)(ph[0] / (sg * sg));
        float sv, dI;
        if (a < X) {
            const float d = (float)(a + b.ox) - (float)ph[2];
            sv = -(d * d) * i2s2;
            dI = d * Iis2;
        } else if (a < X + Y) {
            const float d = (float)(a - X + b.oy) - (float)ph[3];
            sv = -(d * d) * i2s2;
            dI = d * Iis2;
        } else {
            const float dz = (float)(a - X - Y + b.oz) - (float)ph[4];
            const float dzs = dz / zF;
            sv = -(dzs * dzs) * i2s2;
            dI = dz * Iis2 / zdiv;
        }
        ax[a * 8 + j] = expf(sv);
        ax[(nax + a) * 8 + j] = sv;
        ax[(2 * nax + a) * 8 + j] = dI;
    }
    // columns past the spots' basis stay zero
    for (int e = tid; BSTR > NCOL && e < kCorrPix * (BSTR - NCOL); e += 256) {
        const int c = e / (BSTR - NCOL), o = e - c * (BSTR - NCOL);
        Bs[c * BSTR + NCOL + o] = 0.0f;
    }

    // basis role: spot jb, pixel lane cb (pixels p0 + cb of each chunk); the
    // 8 spots of a pixel come from neighbouring lanes, which keeps the basis
    // stores nearly conflict-free.
    const int jb = tid & 7, cb = tid >> 3;
    float Isb = 0.0f, invb = 0.0f, msigb = 0.0f;
    if (jb < nk) {
        const double *ph = b.params + ((size_t)pr.q * K + tk.k0 + jb) * 5;
        Isb = (float)ph[0];
        invb = (float)(1.0 / (ph[1] * 2.5066282746310002));
        msigb = -Isb / (float)ph[1];
    }
    auto basis = [&](int x, int y, int z, float *v) {
        const float *ex = ax + jb;       // exp(-d^2 / 2 sigma^2)
        const float *sx = ex + 8 * nax;  // -d^2 / 2 sigma^2
        const float *dx = sx + 8 * nax;  // d I / sigma^2 (z: over zfactor^2 if exact)
        const int ix = 8 * x, iy = 8 * (X + y), iz = 8 * (X + Y + z);
        const float E = (ex[ix] * ex[iy] * invb) * ex[iz];
        v[0] = Isb * E;
        if (GRAD) {
            const float sv = (sx[ix] + sx[iy]) + sx[iz];
            v[NB > 1 ? 1 : 0] = E;
            v[NB > 2 ? 2 : 0] = E * msigb * (1.0f + 2.0f * sv);
            v[NB > 3 ? 3 : 0] = E * dx[ix];
            v[NB > 4 ? 4 : 0] = E * dx[iy];
            v[NB > 5 ? 5 : 0] = E * dx[iz];
        }
    };
    // Q = sum mu^2 and H_n = sum dmu_n mu over the block (frame independent),
    // accumulated from the chunk loop's basis values: this lane's part is
    // spot jb over the pixels p = cb mod 32.
    // FP32 over four chunks (as the matrix sums), then FP64
    double qacc[NB];
    float q32[NB];
#pragma unroll
    for (int n = 0; n < NB; n++) {
        qacc[n] = 0.0;
        q32[n] = 0.0f;
    }

    const float *Rp = C.R + pr.rslot;
    const int lr = lane & 15, lk = lane >> 4;
    for (int t0 = 0; t0 < T; t0 += kCorrFrames) {
        const int tn = min(kCorrFrames, T - t0);
        v4f acc[2][NCX];
        double a64[2][NCX][4];
#pragma unroll
        for (int f2 = 0; f2 < 2; f2++)
#pragma unroll
            for (int ct = 0; ct < NCX; ct++) {
                acc[f2][ct] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int r = 0; r < 4; r++) a64[f2][ct][r] = 0.0;
            }
        // pixel p = p0 + cb of the basis role, tracked as (x, y, z)
        int bx = (pbeg + cb) % X, by = ((pbeg + cb) / X) % Y, bz = (pbeg + cb) / XY;
        // Residual chunk prefetch: thread owns 4 consecutive pixels
        // (4 (tid & 7)) of rows (tid >> 3) + 32 i; the next chunk's loads are
        // in flight during this chunk's matrix work.
        const int pr4 = (tid & 7) * 4, prow = tid >> 3;
        float4 pre[4];
        auto fetch = [&](int p0) {
            const int p = p0 + pr4;
            if (p0 + kCorrPix <= P) {
                // whole chunk inside the plane: rows past the pass are read
                // from its last row and zeroed
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int r = prow + 32 * i;
                    float4 v;
                    __builtin_memcpy(&v, Rp + (size_t)(t0 + min(r, tn - 1)) * P + p, sizeof v);
                    pre[i] = r < tn ? v : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
                return;
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = prow + 32 * i;
                float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (r < tn) {
                    const float *src = Rp + (size_t)(t0 + r) * P + p;
                    if (p + 3 < P) {
                        __builtin_memcpy(&v, src, sizeof v);  // 4-byte aligned source
                    } else {
                        if (p < P) v.x = src[0];
                        if (p + 1 < P) v.y = src[1];
                        if (p + 2 < P) v.z = src[2];
                    }
                }
                pre[i] = v;
            }
        };
        fetch(pbeg);
        int nch = 0;
        for (int p0 = pbeg; p0 < pend; p0 += kCorrPix, nch++) {
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 4; i++) *(float4 *)(Rs + (prow + 32 * i) * kCorrRow + pr4) = pre[i];
            {
                const int p = p0 + cb;
                float v[NB];
#pragma unroll
                for (int n = 0; n < NB; n++) v[n] = 0.0f;
                // a region given as a box and a mask: pixels outside it
                // carry no spot value (their residuals are zero as well)
                if (jb < nk && p < P && (!FUNC || !b.mask || b.mask[p])) basis(bx, by, bz, v);
#pragma unroll
                for (int n = 0; n < NB; n++) Bs[cb * BSTR + jb * NB + n] = v[n];
                // every pass sees the same values; only the first pass's sums are used
#pragma unroll
                for (int n = 0; n < NB; n++) q32[n] += v[n] * v[0];
                // advance this lane's pixel by kCorrPix
                bx += kCorrPix;
                while (bx >= X) {
                    bx -= X;
                    if (++by == Y) {
                        by = 0;
                        bz++;
                    }
                }
            }
            __syncthreads();
            if (p0 + kCorrPix < pend) fetch(p0 + kCorrPix);
#pragma unroll
            for (int k4 = 0; k4 < kCorrPix; k4 += 4) {
                const int kk = k4 + lk;
                float av[2], bv[NCX];
#pragma unroll
                for (int f2 = 0; f2 < 2; f2++) av[f2] = Rs[((2 * wave + f2) * 16 + lr) * kCorrRow + kk];
#pragma unroll
                for (int ct = 0; ct < NCX; ct++) bv[ct] = Bs[kk * BSTR + ct * 16 + lr];
#pragma unroll
                for (int f2 = 0; f2 < 2; f2++)
#pragma unroll
                    for (int ct = 0; ct < NCX; ct++)
                        acc[f2][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[f2], bv[ct], acc[f2][ct], 0, 0, 0);
            }
            if ((nch & 3) == 3 || p0 + kCorrPix >= pend) {
#pragma unroll
                for (int f2 = 0; f2 < 2; f2++)
#pragma unroll
                    for (int ct = 0; ct < NCX; ct++) {
#pragma unroll
                        for (int r = 0; r < 4; r++) a64[f2][ct][r] += (double)acc[f2][ct][r];
                        acc[f2][ct] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
                    }
#pragma unroll
                for (int n = 0; n < NB; n++) {
                    qacc[n] += (double)q32[n];
                    q32[n] = 0.0f;
                }
            }
        }
        __syncthreads();
        if (t0 == 0) {
            // lanes l, l ^ 8, l ^ 16, l ^ 32 share a spot; then the 4 waves in order
#pragma unroll
            for (int n = 0; n < NB; n++) {
                double v = qacc[n];
                v += __shfl_xor(v, 8);
                v += __shfl_xor(v, 16);
                v += __shfl_xor(v, 32);
                if (lane < kCorrSpots) qw[(wave * kCorrSpots + lane) * 6 + n] = v;
            }
            __syncthreads();
            if (tid < kCorrSpots * NB) {
                const int j = tid / NB, n = tid - j * NB;
                const double *w = qw + j * 6 + n;
                qh[j * 6 + n] = ((w[0] + w[kCorrSpots * 6]) + w[2 * kCorrSpots * 6]) + w[3 * kCorrSpots * 6];
                if (FUNC && nsplit > 1 && j < nk)
                    C.part_qh[((size_t)split * C.NI + tk.item0 + j) * 6 + n] = qh[j * 6 + n];
            }
            __syncthreads();
        }
#pragma unroll
        for (int f2 = 0; f2 < 2; f2++)
#pragma unroll
            for (int ct = 0; ct < NCX; ct++) {
                const int col = ct * 16 + lr;
                const int j = col / NB, n = col - j * NB;
                if (col >= NCOL || j >= nk) continue;
                const int k = tk.k0 + j;
                const int item = tk.item0 + j;
                if (FUNC && nsplit > 1) {
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int trel = (2 * wave + f2) * 16 + 4 * lk + r;
                        if (trel < tn)
                            C.part[(((size_t)split * C.Tmax + t0 + trel) * C.NI + item) * 6 + n] = a64[f2][ct][r];
                    }
                    continue;
                }
                const uint8_t *stk = b.states + ((size_t)pr.s * K + k) * T;
                const double Q = qh[j * 6], Hn = qh[j * 6 + n];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int trel = (2 * wave + f2) * 16 + 4 * lk + r;
                    if (trel >= tn) continue;
                    const int t = t0 + trel;
                    const bool on = stk[t] == 0;
                    const double v = a64[f2][ct][r];
                    double *m = C.em + ((size_t)t * C.NI + item) * 8;
                    if (n == 0) {
                        const double S0 = C.S0[(size_t)tk.plane * C.Tmax + t];
                        const double c2 = 2.0 * v;
                        m[0] = on ? S0 + (c2 + Q) : S0;
                        m[1] = on ? c2 + Q : c2 - Q;
                    } else {
                        m[1 + n] = on ? v : v - Hn;
                    }
                }
            }
    }
}

template __global__ void b3_corr_mfma<false, false, 0>(CorrParams);
template __global__ void b3_corr_mfma<true, false, 0>(CorrParams);
template __global__ void b3_corr_mfma<true, true, 0>(CorrParams);
template __global__ void b3_corr_mfma<true, true, 1>(CorrParams);

// b3_corr_reduce: the pixel-split correlation sums of one (frame, item),
// added over the splits in order in FP64, then the on/off forms of
// b3_corr_mfma's epilogue (Soff = S0 + [on](2C + Q), D = 2C +- Q,
// diff_n = G_n - [off] H_n).
__global__ __launch_bounds__(256) void b3_corr_reduce(CorrReduceParams R) {
    // 8 lanes per (frame, item, field n): lane u adds splits u, u + 8, ...
    // in order, then the 8 partial sums are combined by a fixed butterfly
    // (a fixed order: the result does not depend on the launch).  Field 0
    // gives Soff and D, field n >= 1 gives diff_{n-1}.
    const int g = blockIdx.x * 256 + threadIdx.x;
    const int i = g >> 3, u = g & 7;
    const bool live = i < R.NI * R.T * 6;
    const int ii = live ? i : 0;
    const int n = ii % 6, ti = ii / 6;
    const int t = ti / R.NI, item = ti - t * R.NI;
    const double *__restrict__ pv = R.part + ((size_t)t * R.NI + item) * 6 + n;
    const double *__restrict__ pq = R.part_qh + (size_t)item * 6 + n;
    const size_t sv = (size_t)R.Tmax * R.NI * 6, sq = (size_t)R.NI * 6;
    double v = 0.0, qh = 0.0;
    if (live) {
        for (int sp = u; sp < R.nsplit; sp += 8) {
            v += pv[sp * sv];
            qh += pq[sp * sq];
        }
    }
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
        v += __shfl_xor(v, o);
        qh += __shfl_xor(qh, o);
    }
    if (!live || u != 0) return;
    const Item it = R.items[item];
    const DevBlock &b = *R.block;
    const bool on = b.states[((size_t)it.s * b.K + it.k) * b.T + t] == 0;
    double *m = R.em + ((size_t)t * R.NI + item) * 8;
    if (n == 0) {
        const double S0 = R.S0[(size_t)it.s * R.Tmax + t];
        const double c2 = 2.0 * v;
        m[0] = on ? S0 + (c2 + qh) : S0;
        m[1] = on ? c2 + qh : c2 - qh;
    } else {
        m[1 + n] = on ? v : v - qh;
    }
}

// ---------------------------------------------------------------------------
// Kernel D (b3_hmm): one lane per item runs the FP64 HMM of kernel B from
// the per-frame sums em (forward with checkpoints, logL; for MODE_GRAD the
// adjoint over segments recomputed from the checkpoints, held in ck; for
// MODE_DELTA the forward messages as kernel B writes them).
template <int MODE>
__global__ __launch_bounds__(256) void b3_hmm(HmmParams H) {
    constexpr bool GRAD = MODE == MODE_GRAD;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= H.n) return;
    const int li = H.lo + i;
    const Item it = H.items[li];
    const DevBlock &b = H.blocks[it.block];
    const int T = b.T;
    const DevModel &md = *H.model;
    Hmm h;
    for (int u = 0; u < 9; u++) h.A[u] = md.A[u];
    for (int u = 0; u < 3; u++) h.pi[u] = md.pi[u];
    const double i2v = 0.5 / md.variance;
    const double cst = (double)b.npix * log(2.0 * M_PI * md.variance) / 2.0;
    const int ckpt = H.ckpt > 0 ? min(H.ckpt, T) : (int)ceil(sqrt((double)T));
    const int nseg = (T + ckpt - 1) / ckpt;
    const size_t NI = (size_t)H.NI;
    // Each frame's sums are loaded one frame ahead of the recursion that
    // needs them (the recursion is a serial chain per lane, so a load issued
    // in the iteration that uses it would expose its whole latency).
    double *__restrict__ ck = H.ck + li;               // slot u at ck[u * NI]
    const double *__restrict__ emb = H.em + li * 8;    // frame t at emb[t * es]
    const size_t es = NI * 8;
    const double ninf = -__builtin_inf();
    double am[3] = {0.0, 0.0, 0.0}, ae[3] = {ninf, ninf, ninf};
    double L = 0.0;
    double n0 = emb[0], n1 = emb[1];
    for (int t = 0; t < T; t++) {
        const double c0 = n0, c1 = n1;
        if (t + 1 < T) {
            n0 = emb[(t + 1) * es];
            n1 = emb[(t + 1) * es + 1];
        }
        const double e1 = -c0 * i2v - cst;
        const double dlt = c1 * i2v;
        if (GRAD && t % ckpt == 0) {
            const int u = 6 * (t / ckpt);
#pragma unroll
            for (int j = 0; j < 3; j++) {
                ck[(u + j) * NI] = am[j];
                ck[(u + 3 + j) * NI] = ae[j];
            }
        }
        double pm[3], pe[3], cm, ce;
        xpred(h, am, ae, t == 0, pm, pe);
        if (MODE == MODE_DELTA) {
            // as kernel B: delta_t(j) = ln pred_j + B(j, t) + logL_{t-1}
            double *dout = H.delta + ((size_t)li * T + t) * 3;
            const double Bon = e1 + dlt;
            dout[0] = (pm[0] > 0 ? log(pm[0]) + pe[0] * M_LN2 : -1e100) + Bon + L;
            dout[1] = (pm[1] > 0 ? log(pm[1]) + pe[1] * M_LN2 : -1e100) + e1 + L;
            dout[2] = (pm[2] > 0 ? log(pm[2]) + pe[2] * M_LN2 : -1e100) + e1 + L;
        }
        L += xupdate(pm, pe, xemis(dlt), am, ae, cm, ce) + e1 + fmax(dlt, 0.0);
    }
    const size_t oi = ((size_t)it.q * b.K + it.k) * b.S + it.s;
    b.logL[oi] = L;
    if (!GRAD) return;
    double g[5] = {0, 0, 0, 0, 0};
    double bm[3] = {0.5, 0.5, 0.5}, be[3] = {1.0, 1.0, 1.0};  // beta_{T-1} = 1
    const int sa0 = 6 * nseg, sc0 = sa0 + 2 * ckpt;  // slots: checkpoints | alpha_t(on) | c_t, (m, e) each
    Emis bn = xemis(0.0);
    double cnm = 1.0, cne = 0.0;
    for (int sgi = nseg - 1; sgi >= 0; sgi--) {
        const int t0 = sgi * ckpt, t1 = min(T, t0 + ckpt);
        double arm[3], are[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            arm[j] = ck[(6 * sgi + j) * NI];
            are[j] = ck[(6 * sgi + 3 + j) * NI];
        }
        double p1 = emb[t0 * es + 1];
        for (int t = t0; t < t1; t++) {
            const double c1 = p1;
            if (t + 1 < t1) p1 = emb[(t + 1) * es + 1];
            double cm, ce;
            xfstep(h, arm, are, xemis(c1 * i2v), t == 0, cm, ce);
            ck[(sa0 + 2 * (t - t0)) * NI] = arm[0];
            ck[(sa0 + 2 * (t - t0) + 1) * NI] = are[0];
            ck[(sc0 + 2 * (t - t0)) * NI] = cm;
            ck[(sc0 + 2 * (t - t0) + 1) * NI] = ce;
        }
        // frame t's alpha_t(on), c_t and sums, loaded during frame t + 1
        auto fetch = [&](int t, double (&v)[10]) {
            const double *m8 = emb + t * es;
#pragma unroll
            for (int n = 0; n < 6; n++) v[n] = m8[1 + n];
            v[6] = ck[(sa0 + 2 * (t - t0)) * NI];
            v[7] = ck[(sa0 + 2 * (t - t0) + 1) * NI];
            v[8] = ck[(sc0 + 2 * (t - t0)) * NI];
            v[9] = ck[(sc0 + 2 * (t - t0) + 1) * NI];
        };
        double nx[10];
        fetch(t1 - 1, nx);
        for (int t = t1 - 1; t >= t0; t--) {
            double cur[10];
#pragma unroll
            for (int n = 0; n < 10; n++) cur[n] = nx[n];
            if (t - 1 >= t0) fetch(t - 1, nx);
            if (t < T - 1) xbeta(h, bm, be, bn, cnm, cne);
            const double gam = xvalue(cur[6] * bm[0], cur[7] + be[0]);
            for (int n = 0; n < 5; n++) g[n] += gam * cur[1 + n];
            bn = xemis(cur[0] * i2v);
            cnm = cur[8];
            cne = cur[9];
        }
    }
    const double iv = 1.0 / md.variance;
    for (int n = 0; n < 5; n++) b.grad[oi * 5 + n] = g[n] * iv;
}

template __global__ void b3_hmm<MODE_LOGL>(HmmParams);
template __global__ void b3_hmm<MODE_GRAD>(HmmParams);
template __global__ void b3_hmm<MODE_DELTA>(HmmParams);

// ---------------------------------------------------------------------------
// Kernel D' (b3_hmm_wave): the HMM of kernel D with one wave per item, for
// launches with few items, where one lane per item leaves the GPU idle
// behind a serial chain of T frames.  Lane l takes frames [lF, lF + nf),
// F = ceil(T / 64) <= kWaveFrames.  With the transfer matrix M_t of frame t
// (transfer_matrix: M_t[j][i] = b_t(j) A[i][j], b_t(j) pi_j in frame 0) the
// forward message is alpha_t = M_t ... M_0 e_0 and the backward one
// beta_t^T = 1^T M_{T-1} ... M_{t+1}.  The wave forms the prefix and suffix
// products of the lanes' matrices (6 exchanges each); each lane then runs
// its own frames from the product that enters them and the one that leaves
// them.  Nothing is normalised (the exponents keep the scale), and with the
// emissions relative to e^max(dlt, 0) as in kernel D:
//     logL = ln Z + K_{T-1},   Z = 1^T alpha_{T-1},
//     d logL / d phi = sum_t alpha_t(on) beta_t(on) / Z diff_t / var,
//     delta_t(j) = ln (A^T alpha_{t-1})_j + K_{t-1} + B(j, t)
// where K_t is the running sum of e1 + max(dlt, 0) over frames 0..t.
template <int MODE>
__global__ __launch_bounds__(256) void b3_hmm_wave(HmmParams H) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= H.n) return;  // whole waves
    const int li = H.lo + i;
    const Item it = H.items[li];
    const DevBlock &b = H.blocks[it.block];
    const int T = b.T;
    const DevModel &md = *H.model;
    double A[9], pi[3];
#pragma unroll
    for (int u = 0; u < 9; u++) A[u] = md.A[u];
#pragma unroll
    for (int u = 0; u < 3; u++) pi[u] = md.pi[u];
    const double i2v = 0.5 / md.variance;
    const double cst = (double)b.npix * log(2.0 * M_PI * md.variance) / 2.0;
    const int F = (T + 63) >> 6;             // at most kWaveFrames (the host checks T)
    const int tb = lane * F;                 // the lane's first frame
    const int nf = max(0, min(F, T - tb));   // and its number of frames
    const double *__restrict__ emb = H.em + (size_t)li * 8;  // frame t at emb[t * es]
    const size_t es = (size_t)H.NI * 8;
    const double ninf = -__builtin_inf();

    double dlt[kWaveFrames], e1[kWaveFrames];
    double ks = 0.0;  // the lane's sum of e1 + max(dlt, 0)
#pragma unroll
    for (int f = 0; f < kWaveFrames; f++) {
        dlt[f] = 0.0;
        e1[f] = 0.0;
        if (f < nf) {
            const double *m8 = emb + (size_t)(tb + f) * es;
            e1[f] = -m8[0] * i2v - cst;
            dlt[f] = m8[1] * i2v;
            ks += e1[f] + fmax(dlt[f], 0.0);
        }
    }
    // the lane's product L_l = M_{tb+nf-1} ... M_tb (the identity without frames)
    double Lm[9], Le[9];
    xident(Lm, Le);
#pragma unroll
    for (int f = 0; f < kWaveFrames; f++)
        if (f < nf) {
            double Mm[9], Me[9];
            transfer_matrix(dlt[f], tb + f == 0, A, pi, Mm, Me);
            if (f == 0) {
#pragma unroll
                for (int q = 0; q < 9; q++) {
                    Lm[q] = Mm[q];
                    Le[q] = Me[q];
                }
            } else {
                double Rm[9], Re[9];
                xmatmul_t<false>(Mm, Me, Lm, Le, Rm, Re);
#pragma unroll
                for (int q = 0; q < 9; q++) {
                    Lm[q] = Rm[q];
                    Le[q] = Re[q];
                }
            }
        }
    // alpha entering the lane (column 0 of the previous lane's prefix, e_0 in
    // lane 0) and Z (column 0 of the last lane's prefix, summed)
    double am[3], ae[3], zm, ze;
    {
        double Pm[9], Pe[9];
#pragma unroll
        for (int q = 0; q < 9; q++) {
            Pm[q] = Lm[q];
            Pe[q] = Le[q];
        }
        prefix_products<false>(Pm, Pe, lane);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            am[j] = __shfl_up(Pm[3 * j], 1);
            ae[j] = __shfl_up(Pe[3 * j], 1);
        }
        if (lane == 0) {
            am[0] = 0.5; ae[0] = 1.0;
            am[1] = am[2] = 0.0;
            ae[1] = ae[2] = ninf;
        }
        xsum3n(read_lane(Pm[0], 63), read_lane(Pe[0], 63), read_lane(Pm[3], 63), read_lane(Pe[3], 63),
               read_lane(Pm[6], 63), read_lane(Pe[6], 63), zm, ze);
    }
    // K_{T-1} over the wave; K before the lane's first frame
    double kin = ks;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const double v = __shfl_up(kin, d);
        if (lane >= d) kin += v;
    }
    const double Ktot = read_lane(kin, 63);
    kin = __shfl_up(kin, 1);
    if (lane == 0) kin = 0.0;
    const size_t oi = ((size_t)it.q * b.K + it.k) * b.S + it.s;
    if (lane == 0) b.logL[oi] = zm > 0.0 ? log(zm) + ze * M_LN2 + Ktot : ninf;
    if (MODE == MODE_LOGL) return;

    // forward over the lane's frames: alpha_t(on) for the adjoint, or delta
    double a0m[kWaveFrames], a0e[kWaveFrames];
    double kr = kin;  // K_{t-1}
#pragma unroll
    for (int f = 0; f < kWaveFrames; f++) {
        a0m[f] = 0.0;
        a0e[f] = ninf;
        if (f < nf) {
            const int t = tb + f;
            double pm[3], pe[3];  // A^T alpha_{t-1} (pi in frame 0)
            if (t == 0) {
#pragma unroll
                for (int j = 0; j < 3; j++) xsplit(pi[j], 0.0, pm[j], pe[j]);
            } else {
#pragma unroll
                for (int j = 0; j < 3; j++)
                    xsum3(am[0] * A[j], ae[0], am[1] * A[3 + j], ae[1], am[2] * A[6 + j], ae[2], pm[j], pe[j]);
            }
            if (MODE == MODE_DELTA) {
                double *dout = H.delta + ((size_t)li * T + t) * 3;
                const double Bon = e1[f] + dlt[f];
                dout[0] = (pm[0] > 0 ? log(pm[0]) + pe[0] * M_LN2 : -1e100) + Bon + kr;
                dout[1] = (pm[1] > 0 ? log(pm[1]) + pe[1] * M_LN2 : -1e100) + e1[f] + kr;
                dout[2] = (pm[2] > 0 ? log(pm[2]) + pe[2] * M_LN2 : -1e100) + e1[f] + kr;
                kr += e1[f] + fmax(dlt[f], 0.0);
            }
            const Emis bt = xemis(dlt[f]);
            xsplit(pm[0] * bt.onm, pe[0] + bt.one, am[0], ae[0]);
            xsplit(pm[1] * bt.offm, pe[1] + bt.offe, am[1], ae[1]);
            xsplit(pm[2] * bt.offm, pe[2] + bt.offe, am[2], ae[2]);
            a0m[f] = am[0];
            a0e[f] = ae[0];
        }
    }
    if (MODE == MODE_DELTA) return;

    // beta leaving the lane: 1^T (L_63 ... L_{l+1}), the column sums of the
    // next lane's suffix product (1 in the last lane)
    double bm[3], be[3];
    {
        suffix_products(Lm, Le, lane);
        double wm[3], we[3];
#pragma unroll
        for (int c = 0; c < 3; c++) xsum3n(Lm[c], Le[c], Lm[3 + c], Le[3 + c], Lm[6 + c], Le[6 + c], wm[c], we[c]);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            bm[c] = __shfl_down(wm[c], 1);
            be[c] = __shfl_down(we[c], 1);
        }
        if (lane == kWave - 1) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                bm[c] = 0.5;
                be[c] = 1.0;
            }
        }
    }
    const double izm = zm > 0.0 ? 1.0 / zm : 0.0;
    double g[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int f = kWaveFrames - 1; f >= 0; f--)
        if (f < nf) {
            const int t = tb + f;
            const double gam = xvalue(a0m[f] * bm[0] * izm, a0e[f] + be[0] - ze);
            const double *m8 = emb + (size_t)t * es;
#pragma unroll
            for (int n = 0; n < 5; n++) g[n] += gam * m8[2 + n];
            if (f > 0) {  // beta_{t-1} = M_t^T beta_t
                const Emis bt = xemis(dlt[f]);
                const double wm0 = bt.onm * bm[0], we0 = bt.one + be[0];
                const double wm1 = bt.offm * bm[1], we1 = bt.offe + be[1];
                const double wm2 = bt.offm * bm[2], we2 = bt.offe + be[2];
#pragma unroll
                for (int c = 0; c < 3; c++)
                    xsum3(A[3 * c] * wm0, we0, A[3 * c + 1] * wm1, we1, A[3 * c + 2] * wm2, we2, bm[c], be[c]);
            }
        }
    const double iv = 1.0 / md.variance;
#pragma unroll
    for (int n = 0; n < 5; n++) {
        const double v = wave_sum(g[n]);
        if (lane == 0) b.grad[oi * 5 + n] = v * iv;
    }
}

template __global__ void b3_hmm_wave<MODE_LOGL>(HmmParams);
template __global__ void b3_hmm_wave<MODE_GRAD>(HmmParams);
template __global__ void b3_hmm_wave<MODE_DELTA>(HmmParams);

// ---------------------------------------------------------------------------
// FFBS Hessian, sampled_background_spot_hessian_ffbs (multispot5.cc:902-982).
//
// b3_hess_basis: one thread per pixel of the region's box evaluates
// spot_shape_hess_position (storm.h:86-150) in FP64 and stores the value,
// the upper Hessian and the derivative in FP32; per workgroup it also sums
// d_i d_j (i <= j) in FP64 (the frame-independent part of hess_t).
// z_exact = 0 keeps the reference's z terms (x[2] - phi[4] without 1/zF^2,
// as the derivative) and leaves [4][3], which storm.h:143 never sets, to the
// host (as 0); z_exact = 1 gives the exact second derivatives.
__global__ __launch_bounds__(256) void b3_hess_basis(HessParams Hp) {
    __shared__ double red[4][kHessUpper];
    const DevBlock &b = *Hp.block;
    const DevModel &md = *Hp.model;
    const int X = b.X, Y = b.Y, XY = X * Y, P = XY * b.Z;
    const int p = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double dd[kHessUpper];
#pragma unroll
    for (int k = 0; k < kHessUpper; k++) dd[k] = 0.0;
    if (p < P) {
        float *out = Hp.basis + (size_t)p * kHessBasis;
        float v[kHessBasis];
#pragma unroll
        for (int k = 0; k < kHessBasis; k++) v[k] = 0.0f;
        if (!b.mask || b.mask[p]) {
            const double *ph = b.params;
            const int z = p / XY, y = (p - z * XY) / X, x = p - z * XY - y * X;
            const double zF = md.zfactor;
            const double d0 = (double)(x + b.ox) - ph[2], d1 = (double)(y + b.oy) - ph[3];
            const double d2 = (double)(z + b.oz) - ph[4];
            const double d2s = d2 / zF;
            const double I = ph[0], sg = ph[1];
            const double s = -(d0 * d0 + d1 * d1 + d2s * d2s) / (2 * sg * sg);
            const double r2p = 2.5066282746310002;
            const double es = exp(s);
            const double w[3] = {d0, d1, md.z_exact ? d2 / (zF * zF) : d2};
            const double c[3] = {1.0, 1.0, md.z_exact ? 1.0 / (zF * zF) : 1.0};
            const double sg2 = sg * sg, sg3 = sg2 * sg;
            double d[5], h[kHessUpper];
            const double f = es / (sg * r2p);
            d[0] = f;
            d[1] = f * (-I * (1 + 2 * s) / sg);
            for (int i = 0; i < 3; i++) d[2 + i] = f * (w[i] * (I / sg2));
            h[0] = 0.0;
            h[1] = -es * (1 + 2 * s) / (sg2 * r2p);
            for (int i = 0; i < 3; i++) h[2 + i] = es * w[i] / (sg3 * r2p);
            h[5] = 2 * I * es * (1 + 5 * s + 2 * s * s) / (sg3 * r2p);
            for (int i = 0; i < 3; i++) h[6 + i] = -I * es * (3 + 2 * s) * w[i] / (sg3 * sg * r2p);
            // (2,2) (2,3) (2,4) (3,3) (3,4) (4,4)
            int k = 9;
            for (int i = 0; i < 3; i++)
                for (int j = i; j < 3; j++)
                    h[k++] = I * es * (w[i] * w[j] - (i == j ? sg2 * c[i] : 0.0)) / (r2p * sg3 * sg2);
            v[0] = (float)(es * I / (sg * r2p));
            for (int k2 = 0; k2 < kHessUpper; k2++) v[1 + k2] = (float)h[k2];
            for (int i = 0; i < 5; i++) v[1 + kHessUpper + i] = (float)d[i];
            int u = 0;
            for (int i = 0; i < 5; i++)
                for (int j = i; j < 5; j++) dd[u++] = d[i] * d[j];
        }
#pragma unroll
        for (int k = 0; k < kHessBasis; k += 4)
            *(float4 *)(out + k) = make_float4(v[k], v[k + 1], v[k + 2], v[k + 3]);
    }
#pragma unroll
    for (int k = 0; k < kHessUpper; k++) {
        double a = dd[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
        if (lane == 0) red[wave][k] = a;
    }
    __syncthreads();
    if (threadIdx.x < kHessUpper) {
        const int k = threadIdx.x;
        Hp.fixed[(size_t)blockIdx.x * kHessUpper + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
    }
}

// b3_hess_parts: workgroup (s, 4 frames) sums, over the region's pixels,
// e H_k (15 upper entries) and e d_i (5) with e = R - mu for its frames
// (hess_handler's per-frame parts, multispot5.cc:839-855, without the
// frame-independent d d^T).  Each thread sums its pixels in FP32, then the
// 256 partials are added in FP64 (lanes by exchange, then the 4 waves in
// order).
__global__ __launch_bounds__(256) void b3_hess_parts(HessParams Hp) {
    __shared__ double red[4][kHessFrames * kHessParts];
    const DevBlock &b = *Hp.block;
    const int P = b.X * b.Y * b.Z, T = Hp.T;
    const int ng = (T + kHessFrames - 1) / kHessFrames;
    const int s = blockIdx.x / ng, t0 = (blockIdx.x - s * ng) * kHessFrames;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nf = min(kHessFrames, T - t0);
    const float *Rs = Hp.R + ((size_t)s * T + t0) * P;
    float acc[kHessFrames][kHessParts];
#pragma unroll
    for (int f = 0; f < kHessFrames; f++)
#pragma unroll
        for (int k = 0; k < kHessParts; k++) acc[f][k] = 0.0f;
    for (int p = threadIdx.x; p < P; p += 256) {
        const float4 *bp = (const float4 *)(Hp.basis + (size_t)p * kHessBasis);
        float v[kHessBasis];
#pragma unroll
        for (int k = 0; k < kHessBasis / 4; k++) {
            const float4 q = bp[k];
            v[4 * k] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
        }
#pragma unroll
        for (int f = 0; f < kHessFrames; f++) {
            if (f < nf) {
                const float e = Rs[(size_t)f * P + p] - v[0];
#pragma unroll
                for (int k = 0; k < kHessParts; k++) acc[f][k] += e * v[1 + k];
            }
        }
    }
#pragma unroll
    for (int f = 0; f < kHessFrames; f++)
#pragma unroll
        for (int k = 0; k < kHessParts; k++) {
            double a = (double)acc[f][k];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
            if (lane == 0) red[wave][f * kHessParts + k] = a;
        }
    __syncthreads();
    for (int i = threadIdx.x; i < kHessFrames * kHessParts; i += 256) {
        const int f = i / kHessParts, k = i - f * kHessParts;
        if (f < nf)
            Hp.parts[((size_t)s * T + t0 + f) * kHessParts + k] = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
    }
}

// ---------------------------------------------------------------------------
// Backward sampling of forward filtering / backward sampling, one lane per
// chain (forward_algorithm.h:502-566): the last frame's state is drawn from
// its forward message, then each earlier frame's from its message plus
// ln A[j][next]; a draw picks the first state whose running sum of
// exp(a - max) reaches u * sum.  Draw k of a chain is u[k] (u[0]: last frame).
struct LogA {
    double v[9];  // ln A[i][j], ln 0 = -1e100
};

__device__ __forceinline__ int sample_unscaled_log3(double a0, double a1, double a2, double u) {
    const double hi = fmax(a0, fmax(a1, a2));
    const double e0 = exp(a0 - hi), e1 = exp(a1 - hi), e2 = exp(a2 - hi);
    double sum = 0.0;
    sum += e0;
    sum += e1;
    sum += e2;
    const double choice = u * sum;
    double total = e0;
    if (choice <= total) return 0;
    total += e1;
    if (choice <= total) return 1;
    return 2;  // either the third running sum reaches choice or the last element is returned
}

// Chain c samples from the messages delta[c / per] (per chains share one
// set of messages: repeated backward samplings of one forward pass).
__global__ void b3_backward_sample(const double *__restrict__ delta, int n, int per, int T, LogA lA,
                                   const double *__restrict__ u, uint8_t *__restrict__ states) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double *d = delta + (size_t)(c / per) * T * 3;
    const double *uc = u + (size_t)c * T;
    uint8_t *sc = states + (size_t)c * T;
    int k = 0;
    int next = sample_unscaled_log3(d[3 * (T - 1)], d[3 * (T - 1) + 1], d[3 * (T - 1) + 2], uc[k++]);
    sc[T - 1] = (uint8_t)next;
    for (int t = T - 2; t >= 0; t--) {
        next = sample_unscaled_log3(d[3 * t] + lA.v[next], d[3 * t + 1] + lA.v[3 + next], d[3 * t + 2] + lA.v[6 + next],
                                    uc[k++]);
        sc[t] = (uint8_t)next;
    }
}

template __global__ void b3_eval<MODE_LOGL>(EvalParams);
template __global__ void b3_eval<MODE_GRAD>(EvalParams);
template __global__ void b3_eval<MODE_DELTA>(EvalParams);

}  // namespace rfba3b
