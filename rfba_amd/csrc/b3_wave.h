// b3_wave.h -- wave-wide helpers of the 3B kernels (one 64-lane wave per
// chain or item): sums and lane reads, and 3 x 3 transfer-matrix products
// in extended range with prefix and suffix products across the wave.
#pragma once
#include "b3_internal.h"

namespace rfba3b {

constexpr int kWave = 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// The value of lane l for the whole wave (v_readlane); l must be uniform.
__device__ __forceinline__ double read_lane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// The identity in extended range (1 = 0.5 2^1, 0 = 0 2^-inf).
__device__ __forceinline__ void xident(double (&Mm)[9], double (&Me)[9]) {
#pragma unroll
    for (int q = 0; q < 9; q++) {
        Mm[q] = q % 4 == 0 ? 0.5 : 0.0;
        Me[q] = q % 4 == 0 ? 1.0 : -__builtin_inf();
    }
}

// The transfer matrix of a frame in extended range: M[j][i] = b_j A[i][j]
// (b_j pi_j in frame 0), b_on = e^min(x,0), b_off = e^min(-x,0) for the
// frame's x = B_on - B_off (states 1 and 2 are off).
__device__ __forceinline__ void transfer_matrix(double x, bool first, const double *A, const double *pi,
                                                double (&Mm)[9], double (&Me)[9]) {
    double bm[2], be[2];
    xexp(fmin(x, 0.0), bm[0], be[0]);   // b_on
    xexp(fmin(-x, 0.0), bm[1], be[1]);  // b_off
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++)
            xsplit((first ? pi[j] : A[3 * i + j]) * bm[j == 0 ? 0 : 1], be[j == 0 ? 0 : 1], Mm[3 * j + i],
                   Me[3 * j + i]);
}

// R = S Q for 3 x 3 matrices in extended range (row-major, entry (j, i) at
// 3 j + i); SHIFT: shifted so that the largest entry has exponent 0..1,
// else the exponents keep the product's scale.
template <bool SHIFT>
__device__ __forceinline__ void xmatmul_t(const double (&Sm)[9], const double (&Se)[9], const double (&Qm)[9],
                                          const double (&Qe)[9], double (&Rm)[9], double (&Re)[9]) {
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++)
            xsum3n(Sm[3 * j] * Qm[i], Se[3 * j] + Qe[i], Sm[3 * j + 1] * Qm[3 + i], Se[3 * j + 1] + Qe[3 + i],
                   Sm[3 * j + 2] * Qm[6 + i], Se[3 * j + 2] + Qe[6 + i], Rm[3 * j + i], Re[3 * j + i]);
    if (SHIFT) xshift(Re);
}

__device__ __forceinline__ void xmatmul(const double (&Sm)[9], const double (&Se)[9], const double (&Qm)[9],
                                        const double (&Qe)[9], double (&Rm)[9], double (&Re)[9]) {
    xmatmul_t<true>(Sm, Se, Qm, Qe, Rm, Re);
}

// w = M v in extended range, shifted so that its largest entry has
// exponent 0..1.
__device__ __forceinline__ void xmatvec(const double (&Mm)[9], const double (&Me)[9], const double (&vm)[3],
                                        const double (&ve)[3], double (&wm)[3], double (&we)[3]) {
#pragma unroll
    for (int j = 0; j < 3; j++)
        xsum3n(Mm[3 * j] * vm[0], Me[3 * j] + ve[0], Mm[3 * j + 1] * vm[1], Me[3 * j + 1] + ve[1],
               Mm[3 * j + 2] * vm[2], Me[3 * j + 2] + ve[2], wm[j], we[j]);
    xshift(we);
}

// Prefix products across the wave, in extended range: on entry (Mm, Me) is
// lane l's matrix L_l; on return it is L_l L_{l-1} ... L_0.
template <bool SHIFT>
__device__ __forceinline__ void prefix_products(double (&Mm)[9], double (&Me)[9], int lane) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        double Qm[9], Qe[9];
#pragma unroll
        for (int q = 0; q < 9; q++) {
            Qm[q] = __shfl_up(Mm[q], d);
            Qe[q] = __shfl_up(Me[q], d);
        }
        if (lane >= d) {
            double Rm[9], Re[9];
            xmatmul_t<SHIFT>(Mm, Me, Qm, Qe, Rm, Re);
#pragma unroll
            for (int q = 0; q < 9; q++) {
                Mm[q] = Rm[q];
                Me[q] = Re[q];
            }
        }
    }
}

// Suffix products across the wave, unshifted: on entry (Mm, Me) is lane
// l's matrix L_l; on return it is L_63 L_62 ... L_l.
__device__ __forceinline__ void suffix_products(double (&Mm)[9], double (&Me)[9], int lane) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        double Qm[9], Qe[9];
#pragma unroll
        for (int q = 0; q < 9; q++) {
            Qm[q] = __shfl_down(Mm[q], d);
            Qe[q] = __shfl_down(Me[q], d);
        }
        if (lane + d < kWave) {
            double Rm[9], Re[9];
            xmatmul_t<false>(Qm, Qe, Mm, Me, Rm, Re);
#pragma unroll
            for (int q = 0; q < 9; q++) {
                Mm[q] = Rm[q];
                Me[q] = Re[q];
            }
        }
    }
}

}  // namespace rfba3b
