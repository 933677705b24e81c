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
        // tables: gxy[K][X+Y] (x | y factors, padded to 4 floats), gz[K][Zp]
        // (z factors, Zp = Z rounded up to 8, zero past Z: 16-byte aligned
        // rows read as two float4 per 8-deep run), amp[K], on[K]
        const int nxy = X + Y, nxyp = (K * nxy + 3) & ~3, Zp = (Z + 7) & ~7;
        float *gxy = lds;
        float *gzt = gxy + nxyp;
        float *amp = gzt + K * Zp;
        int *on = (int *)(amp + K);
        __syncthreads();
        const double zF = mdl->zfactor;
        for (int e = tid; e < K * nxy; e += 256) {
            const int j = e / nxy, a = e - j * nxy;
            const double *ph = b.params + ((size_t)tk.q * K + j) * 5;
            const float sig = (float)ph[1];
            const float d = a < X ? (float)(a + b.ox) - (float)ph[2] : (float)(a - X + b.oy) - (float)ph[3];
            gxy[e] = expf(-(d * d) / (2.0f * sig * sig));
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
        for (int j = tid; j < K; j += 256) {
            const double *ph = b.params + ((size_t)tk.q * K + j) * 5;
            amp[j] = (float)(ph[0] / (ph[1] * 2.5066282746310002));
            on[j] = b.states[((size_t)tk.s * K + j) * T + tk.t] == 0;
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
            for (int j = 0; j < K; j++) {
                if (!on[j]) continue;
                const float *gj = gxy + j * nxy;
                const float a = amp[j] * gj[x] * gj[X + y];
                const float4 *gz = (const float4 *)(gzt + j * Zp + z0);
                const float4 g0 = gz[0], g1 = gz[1];
                m[0] += a * g0.x; m[1] += a * g0.y; m[2] += a * g0.z; m[3] += a * g0.w;
                m[4] += a * g1.x; m[5] += a * g1.y; m[6] += a * g1.z; m[7] += a * g1.w;
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

template <bool GRAD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void b3_corr_mfma(CorrParams C) {
    extern __shared__ double ldsc[];
    constexpr int NB = GRAD ? 6 : 1;
    constexpr int NCOL = kCorrSpots * NB;   // 48 or 8 basis columns
    constexpr int NCT = (NCOL + 15) / 16;   // 16-column tiles
    // basis chunk row stride: a multiple of 16 that is 16 or 48 mod 64, so the
    // four pixel rows a B-operand read touches fall on disjoint bank groups
    constexpr int BSTR = NCT * 16;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // Workgroup b runs on XCD b % 8: give each XCD a contiguous range of
    // tasks so that the spot groups of one plane share that XCD's L2 (the
    // grid is padded to a multiple of 8).
    const int per_xcd = gridDim.x >> 3;
    const int task = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (task >= C.ntasks) return;
    const CorrTask tk = C.tasks[task];
    const PlaneRef pr = C.planes[tk.plane];
    const DevBlock &b = C.blocks[pr.block];
    const int X = b.X, Y = b.Y, Z = b.Z, T = b.T, K = b.K;
    const int XY = X * Y, P = XY * Z, nax = X + Y + Z;
    const int nk = tk.nk;
    const DevModel &md = *C.model;

    double *qh = ldsc;                                   // [8][6] Q, H_0..4
    float *Rs = (float *)(qh + kCorrSpots * 6);          // [kCorrFrames][kCorrRow]
    float *Bs = Rs + kCorrFrames * kCorrRow;             // [kCorrPix][BSTR]
    float *ax = Bs + kCorrPix * BSTR;                    // [8][nax][4]

    const float zF = (float)md.zfactor;
    const float zdiv = md.z_exact ? zF * zF : 1.0f;
    for (int e = tid; e < nk * nax; e += 256) {
        const int j = e / nax, a = e - j * nax;
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
        float *q = ax + 4 * (j * nax + a);
        q[0] = expf(sv);
        q[1] = sv;
        q[2] = dI;
    }
    // columns past the spots' basis stay zero
    for (int e = tid; BSTR > NCOL && e < kCorrPix * (BSTR - NCOL); e += 256) {
        const int c = e / (BSTR - NCOL), o = e - c * (BSTR - NCOL);
        Bs[c * BSTR + NCOL + o] = 0.0f;
    }

    // basis roles: spot jb, pixel lane cb.  The Q/H pre-pass reduces over
    // the 32 lanes of a half-wave (jb = tid >> 5); the chunk loop writes the
    // 8 spots of a pixel from neighbouring lanes (jb = tid & 7), which keeps
    // its LDS stores nearly conflict-free.
    int jb = tid >> 5, cb = tid & 31;
    float Isb = 0.0f, invb = 0.0f, msigb = 0.0f;
    auto spot_consts = [&]() {
        Isb = invb = msigb = 0.0f;
        if (jb < nk) {
            const double *ph = b.params + ((size_t)pr.q * K + tk.k0 + jb) * 5;
            Isb = (float)ph[0];
            invb = (float)(1.0 / (ph[1] * 2.5066282746310002));
            msigb = -Isb / (float)ph[1];
        }
    };
    spot_consts();
    // Q = sum mu^2 and H_n = sum dmu_n mu over the block (frame independent):
    // spot jb, pixels cb, cb + 32, ...
    auto basis = [&](int x, int y, int z, float *v) {
        const float *qx = ax + 4 * (jb * nax + x);
        const float *qy = ax + 4 * (jb * nax + X + y);
        const float *qz = ax + 4 * (jb * nax + X + Y + z);
        const float E = (qx[0] * qy[0] * invb) * qz[0];
        v[0] = Isb * E;
        if (GRAD) {
            const float sv = (qx[1] + qy[1]) + qz[1];
            v[NB > 1 ? 1 : 0] = E;
            v[NB > 2 ? 2 : 0] = E * msigb * (1.0f + 2.0f * sv);
            v[NB > 3 ? 3 : 0] = E * qx[2];
            v[NB > 4 ? 4 : 0] = E * qy[2];
            v[NB > 5 ? 5 : 0] = E * qz[2];
        }
    };
    __syncthreads();  // axis tables
    {
        double qacc[NB];
#pragma unroll
        for (int n = 0; n < NB; n++) qacc[n] = 0.0;
        if (jb < nk)
            for (int p = cb; p < P; p += 32) {
                const int z = p / XY, rr = p - z * XY;
                const int y = rr / X, x = rr - y * X;
                float v[NB];
                basis(x, y, z, v);
#pragma unroll
                for (int n = 0; n < NB; n++) qacc[n] += (double)v[n] * (double)v[0];
            }
#pragma unroll
        for (int n = 0; n < NB; n++) {
            double v = qacc[n];
#pragma unroll
            for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if (cb == 0) qh[jb * 6 + n] = v;
        }
    }

    jb = tid & 7;
    cb = tid >> 3;
    spot_consts();
    const float *Rp = C.R + pr.rslot;
    const int lr = lane & 15, lk = lane >> 4;
    for (int t0 = 0; t0 < T; t0 += kCorrFrames) {
        const int tn = min(kCorrFrames, T - t0);
        v4f acc[2][NCT];
        double a64[2][NCT][4];
#pragma unroll
        for (int f2 = 0; f2 < 2; f2++)
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) {
                acc[f2][ct] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int r = 0; r < 4; r++) a64[f2][ct][r] = 0.0;
            }
        // pixel p = p0 + cb of the basis role, tracked as (x, y, z)
        int bx = cb % X, by = (cb / X) % Y, bz = cb / XY;
        // Residual chunk prefetch: thread owns 4 consecutive pixels
        // (4 (tid & 7)) of rows (tid >> 3) + 32 i; the next chunk's loads are
        // in flight during this chunk's matrix work.
        const int pr4 = (tid & 7) * 4, prow = tid >> 3;
        float4 pre[4];
        auto fetch = [&](int p0) {
            const int p = p0 + pr4;
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
        fetch(0);
        int nch = 0;
        for (int p0 = 0; p0 < P; p0 += kCorrPix, nch++) {
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 4; i++) *(float4 *)(Rs + (prow + 32 * i) * kCorrRow + pr4) = pre[i];
            {
                const int p = p0 + cb;
                float v[NB];
#pragma unroll
                for (int n = 0; n < NB; n++) v[n] = 0.0f;
                if (jb < nk && p < P) basis(bx, by, bz, v);
#pragma unroll
                for (int n = 0; n < NB; n++) Bs[cb * BSTR + jb * NB + n] = v[n];
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
            if (p0 + kCorrPix < P) fetch(p0 + kCorrPix);
#pragma unroll
            for (int k4 = 0; k4 < kCorrPix; k4 += 4) {
                const int kk = k4 + lk;
                float av[2], bv[NCT];
#pragma unroll
                for (int f2 = 0; f2 < 2; f2++) av[f2] = Rs[((2 * wave + f2) * 16 + lr) * kCorrRow + kk];
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) bv[ct] = Bs[kk * BSTR + ct * 16 + lr];
#pragma unroll
                for (int f2 = 0; f2 < 2; f2++)
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++)
                        acc[f2][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[f2], bv[ct], acc[f2][ct], 0, 0, 0);
            }
            if ((nch & 3) == 3 || p0 + kCorrPix >= P) {
#pragma unroll
                for (int f2 = 0; f2 < 2; f2++)
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++) {
#pragma unroll
                        for (int r = 0; r < 4; r++) a64[f2][ct][r] += (double)acc[f2][ct][r];
                        acc[f2][ct] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
                    }
            }
        }
        __syncthreads();
#pragma unroll
        for (int f2 = 0; f2 < 2; f2++)
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) {
                const int col = ct * 16 + lr;
                const int j = col / NB, n = col - j * NB;
                if (col >= NCOL || j >= nk) continue;
                const int k = tk.k0 + j;
                const int item = tk.item0 + j;
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

template __global__ void b3_corr_mfma<false>(CorrParams);
template __global__ void b3_corr_mfma<true>(CorrParams);

// ---------------------------------------------------------------------------
// Kernel D (b3_hmm): one lane per item runs the FP64 HMM of kernel B from
// the per-frame sums em (forward with checkpoints, logL; for GRAD the
// adjoint over segments recomputed from the checkpoints, held in ck).
template <bool GRAD>
__global__ __launch_bounds__(256) void b3_hmm(HmmParams H) {
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
    double *ck = H.ck + li;  // slot u at ck[u * NI]
    auto emr = [&](int t) { return H.em + ((size_t)t * NI + li) * 8; };
    double a[3] = {0, 0, 0};
    double L = 0.0;
    for (int t = 0; t < T; t++) {
        const double *m8 = emr(t);
        const double e1 = -m8[0] * i2v - cst;
        double bon, boff, m;
        emis(m8[1] * i2v, bon, boff, m);
        if (t % ckpt == 0) {
            const int u = 3 * (t / ckpt);
            ck[u * NI] = a[0];
            ck[(u + 1) * NI] = a[1];
            ck[(u + 2) * NI] = a[2];
        }
        L += fstep(h, a, bon, boff, t == 0) + e1 + m;
    }
    const size_t oi = ((size_t)it.q * b.K + it.k) * b.S + it.s;
    b.logL[oi] = L;
    if (!GRAD) return;
    double g[5] = {0, 0, 0, 0, 0};
    double beta[3] = {1.0, 1.0, 1.0};
    const int sa0 = 3 * nseg, sc0 = sa0 + ckpt;  // slots: checkpoints | alpha_t(on) | log c_t
    double c_next = 0.0, bon_n = 0.0, boff_n = 0.0;
    for (int sgi = nseg - 1; sgi >= 0; sgi--) {
        const int t0 = sgi * ckpt, t1 = min(T, t0 + ckpt);
        double ar[3] = {ck[(3 * sgi) * NI], ck[(3 * sgi + 1) * NI], ck[(3 * sgi + 2) * NI]};
        for (int t = t0; t < t1; t++) {
            double bon, boff, m;
            emis(emr(t)[1] * i2v, bon, boff, m);
            ck[(sc0 + t - t0) * NI] = fstep(h, ar, bon, boff, t == 0);
            ck[(sa0 + t - t0) * NI] = ar[0];
        }
        for (int t = t1 - 1; t >= t0; t--) {
            if (t < T - 1) {
                const double ic = exp(-c_next);
                const double w0 = bon_n * beta[0], w1 = boff_n * beta[1], w2 = boff_n * beta[2];
                const double n0 = (h.A[0] * w0 + h.A[1] * w1 + h.A[2] * w2) * ic;
                const double n1 = (h.A[3] * w0 + h.A[4] * w1 + h.A[5] * w2) * ic;
                const double n2 = (h.A[6] * w0 + h.A[7] * w1 + h.A[8] * w2) * ic;
                beta[0] = n0; beta[1] = n1; beta[2] = n2;
            }
            const double *m8 = emr(t);
            const double gam = ck[(sa0 + t - t0) * NI] * beta[0];
            for (int n = 0; n < 5; n++) g[n] += gam * m8[2 + n];
            double m;
            emis(m8[1] * i2v, bon_n, boff_n, m);
            c_next = ck[(sc0 + t - t0) * NI];
        }
    }
    const double iv = 1.0 / md.variance;
    for (int n = 0; n < 5; n++) b.grad[oi * 5 + n] = g[n] * iv;
}

template __global__ void b3_hmm<false>(HmmParams);
template __global__ void b3_hmm<true>(HmmParams);

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

__global__ void b3_backward_sample(const double *__restrict__ delta, int n, int T, LogA lA,
                                   const double *__restrict__ u, uint8_t *__restrict__ states) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double *d = delta + (size_t)c * T * 3;
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
