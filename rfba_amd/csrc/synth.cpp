// synth.cpp -- seeded synthetic fluorescence series (librfba_synth.so).
//
// Workload generator for the SRRF benchmarks and tests: emitters at uniform
// sub-voxel positions, a separable 3D Gaussian PSF evaluated at integer voxel
// coordinates, two-state Markov blinking started from its stationary
// distribution, a constant background and Poisson counts clipped to uint16.
// Every frame plane is drawn from its own counter-seeded stream, so the
// output is bit-identical for any thread count.
//
// Poisson sampling: PTRS, W. Hoermann, "The transformed rejection method for
// generating Poisson random variables", Insurance: Mathematics and Economics
// 12 (1993) 39-45, for lambda >= 10; Knuth's product method below.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto &v : s) v = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {  // xoshiro256**
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double normal() {  // Box-Muller (one value per call)
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

long poisson(Rng &r, double lam) {
    if (lam <= 0) return 0;
    if (lam < 10.0) {
        const double L = std::exp(-lam);
        long k = 0;
        double p = 1.0;
        do {
            k++;
            p *= r.uniform();
        } while (p > L);
        return k - 1;
    }
    const double slam = std::sqrt(lam), loglam = std::log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2);
    for (;;) {
        const double U = r.uniform() - 0.5;
        const double V = r.uniform();
        const double us = 0.5 - std::fabs(U);
        const double k = std::floor((2 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return (long)k;
        if (k < 0 || (us < 0.013 && V > us)) continue;
        if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <=
            -lam + k * loglam - std::lgamma(k + 1))
            return (long)k;
    }
}

template <typename F> void parallel(int n, int threads, F f) {
    threads = std::max(1, std::min(threads, n));
    std::vector<std::thread> th;
    for (int i = 0; i < threads; i++)
        th.emplace_back([=] {
            for (int j = i; j < n; j += threads) f(j);
        });
    for (auto &t : th) t.join();
}

}  // namespace

extern "C" {

// Draw emitter positions and on/off states.  pos: 3*n doubles (x, y, z);
// on: T*n bytes (frame-major).  Initial state from the stationary law
// P(on) = p_off_on / (p_on_off + p_off_on).
int rfba_synth_emitters(int W, int H, int D, int T, int n, uint64_t seed, double p_on_off,
                        double p_off_on, double *pos, uint8_t *on) {
    if (W <= 0 || H <= 0 || D <= 0 || T <= 0 || n < 0) return -1;
    Rng r(seed);
    for (int e = 0; e < n; e++) {
        pos[3 * e] = r.uniform() * W;
        pos[3 * e + 1] = r.uniform() * H;
        pos[3 * e + 2] = r.uniform() * D;
    }
    const double p_stat = p_off_on / (p_on_off + p_off_on);
    for (int e = 0; e < n; e++) on[e] = r.uniform() < p_stat;
    for (int t = 1; t < T; t++)
        for (int e = 0; e < n; e++) {
            const uint8_t prev = on[(size_t)(t - 1) * n + e];
            const double u = r.uniform();
            on[(size_t)t * n + e] = prev ? (u >= p_on_off) : (u < p_off_on);
        }
    return 0;
}

// Expected photon counts lambda[t][z][y][x] (double) or Poisson counts as
// uint16.  The PSF is truncated at 5 sigma per axis.
static void frame_lambda(int W, int H, int D, int n, const double *pos, const uint8_t *on_t, double sxy,
                         double sz, double peak, double bg, std::vector<double> &lam) {
    std::fill(lam.begin(), lam.end(), bg);
    const int rxy = (int)std::ceil(5 * sxy), rz = (int)std::ceil(5 * sz);
    std::vector<double> gx(2 * rxy + 1), gy(2 * rxy + 1), gz(2 * rz + 1);
    for (int e = 0; e < n; e++) {
        if (!on_t[e]) continue;
        const double ex = pos[3 * e], ey = pos[3 * e + 1], ez = pos[3 * e + 2];
        const int cx = (int)std::floor(ex), cy = (int)std::floor(ey), cz = (int)std::floor(ez);
        for (int i = -rxy; i <= rxy; i++) {
            const double dx = (cx + i) - ex, dy = (cy + i) - ey;
            gx[i + rxy] = std::exp(-dx * dx / (2 * sxy * sxy));
            gy[i + rxy] = std::exp(-dy * dy / (2 * sxy * sxy));
        }
        for (int k = -rz; k <= rz; k++) {
            const double dz = (cz + k) - ez;
            gz[k + rz] = std::exp(-dz * dz / (2 * sz * sz));
        }
        for (int k = -rz; k <= rz; k++) {
            const int z = cz + k;
            if (z < 0 || z >= D) continue;
            for (int j = -rxy; j <= rxy; j++) {
                const int y = cy + j;
                if (y < 0 || y >= H) continue;
                const double gyz = peak * gy[j + rxy] * gz[k + rz];
                double *row = lam.data() + ((size_t)z * H + y) * W;
                for (int i = -rxy; i <= rxy; i++) {
                    const int x = cx + i;
                    if (x < 0 || x >= W) continue;
                    row[x] += gyz * gx[i + rxy];
                }
            }
        }
    }
}

int rfba_synth_srrf_u16(int W, int H, int D, int T, int n, uint64_t seed, double sxy, double sz,
                        double peak, double bg, double p_on_off, double p_off_on, uint16_t *out,
                        int threads) {
    if (W <= 0 || H <= 0 || D <= 0 || T <= 0 || n < 0 || !out) return -1;
    std::vector<double> pos(3 * (size_t)n);
    std::vector<uint8_t> on((size_t)T * n);
    if (rfba_synth_emitters(W, H, D, T, n, seed, p_on_off, p_off_on, pos.data(), on.data())) return -1;
    const size_t N = (size_t)W * H * D;
    parallel(T, threads, [&](int t) {
        std::vector<double> lam(N);
        frame_lambda(W, H, D, n, pos.data(), on.data() + (size_t)t * n, sxy, sz, peak, bg, lam);
        for (int z = 0; z < D; z++) {
            uint64_t key = seed ^ 0xA5A5A5A55A5A5A5AULL;
            key += (uint64_t)t * 0x100000001B3ULL + (uint64_t)z * 0x9E3779B97F4A7C15ULL;
            Rng r(splitmix64(key));
            for (size_t i = (size_t)z * W * H; i < (size_t)(z + 1) * W * H; i++) {
                const long c = poisson(r, lam[i]);
                out[(size_t)t * N + i] = (uint16_t)std::min<long>(c, 65535);
            }
        }
    });
    return 0;
}

// Expected-count series (no noise), for closed-form checks.
int rfba_synth_lambda_f64(int W, int H, int D, int T, int n, uint64_t seed, double sxy, double sz,
                          double peak, double bg, double p_on_off, double p_off_on, double *out) {
    std::vector<double> pos(3 * (size_t)n);
    std::vector<uint8_t> on((size_t)T * n);
    if (rfba_synth_emitters(W, H, D, T, n, seed, p_on_off, p_off_on, pos.data(), on.data())) return -1;
    const size_t N = (size_t)W * H * D;
    std::vector<double> lam(N);
    for (int t = 0; t < T; t++) {
        frame_lambda(W, H, D, n, pos.data(), on.data() + (size_t)t * n, sxy, sz, peak, bg, lam);
        std::copy(lam.begin(), lam.end(), out + (size_t)t * N);
    }
    return 0;
}

// One 3B block: K spots with the reference spot shape (storm.h:229-235)
// I/(sigma sqrt(2 pi)) exp(-(dx^2 + dy^2 + (dz/zF)^2)/(2 sigma^2)) at integer
// voxel coordinates, states sampled from the 3-state HMM (A row-major,
// 0 = on), data = b0 + sum of on spots + N(0, var).  data [T][Z][Y][X],
// params [K][5] = (I, sigma, x, y, z), states [K][T].
int rfba_synth_3b_block(int X, int Y, int Z, int T, int K, uint64_t seed, double I, double sigma, double zF,
                        double b0, double var, const double *A, const double *pi, float *data, double *params,
                        uint8_t *states) {
    if (X <= 0 || Y <= 0 || Z <= 0 || T <= 0 || K < 0) return -1;
    Rng r(seed);
    for (int k = 0; k < K; k++) {
        params[5 * k] = I;
        params[5 * k + 1] = sigma;
        params[5 * k + 2] = r.uniform() * (X - 1);
        params[5 * k + 3] = r.uniform() * (Y - 1);
        params[5 * k + 4] = r.uniform() * (Z - 1);
    }
    for (int k = 0; k < K; k++) {
        double u = r.uniform();
        int st = u < pi[0] ? 0 : (u < pi[0] + pi[1] ? 1 : 2);
        states[(size_t)k * T] = (uint8_t)st;
        for (int t = 1; t < T; t++) {
            u = r.uniform();
            const double *row = A + 3 * st;
            st = u < row[0] ? 0 : (u < row[0] + row[1] ? 1 : 2);
            states[(size_t)k * T + t] = (uint8_t)st;
        }
    }
    const size_t P = (size_t)X * Y * Z;
    std::vector<double> mu(P);
    std::vector<double> img(P);
    const double r2pi = std::sqrt(2 * M_PI);
    for (int t = 0; t < T; t++) {
        std::fill(img.begin(), img.end(), b0);
        for (int k = 0; k < K; k++) {
            if (states[(size_t)k * T + t] != 0) continue;
            const double *ph = params + 5 * k;
            for (size_t p = 0; p < P; p++) {
                const int x = (int)(p % X), y = (int)((p / X) % Y), z = (int)(p / ((size_t)X * Y));
                const double dx = x - ph[2], dy = y - ph[3], dz = (z - ph[4]) / zF;
                img[p] += std::exp(-(dx * dx + dy * dy + dz * dz) / (2 * ph[1] * ph[1])) * ph[0] / (ph[1] * r2pi);
            }
        }
        const double sd = std::sqrt(var);
        for (size_t p = 0; p < P; p++) data[(size_t)t * P + p] = (float)(img[p] + sd * r.normal());
    }
    return 0;
}

// Many 3B blocks at once (threads over blocks), each as rfba_synth_3b_block
// but with the spot truncated to |dx|,|dy| <= 5 sigma, |dz| <= 5 sigma zF
// and evaluated from separable per-axis factors.  Block i: dims[3*i..],
// K[i] spots, seed seed0 + i, outputs at data + data_off[i],
// params + 5*spot_off[i], states + state_off[i].
int rfba_synth_3b_blocks(int n, const int *dims, int T, const int *K, uint64_t seed0, double I, double sigma,
                         double zF, double b0, double var, const double *A, const double *pi, float *data,
                         const long long *data_off, double *params, const long long *spot_off, uint8_t *states,
                         const long long *state_off, int threads) {
    parallel(n, threads, [&](int i) {
        const int X = dims[3 * i], Y = dims[3 * i + 1], Z = dims[3 * i + 2], Kb = K[i];
        Rng r(seed0 + (uint64_t)i * 0x9E3779B97F4A7C15ULL);
        double *pp = params + 5 * spot_off[i];
        uint8_t *ss = states + state_off[i];
        for (int k = 0; k < Kb; k++) {
            pp[5 * k] = I;
            pp[5 * k + 1] = sigma;
            pp[5 * k + 2] = r.uniform() * (X - 1);
            pp[5 * k + 3] = r.uniform() * (Y - 1);
            pp[5 * k + 4] = r.uniform() * (Z - 1);
        }
        for (int k = 0; k < Kb; k++) {
            double u = r.uniform();
            int st = u < pi[0] ? 0 : (u < pi[0] + pi[1] ? 1 : 2);
            ss[(size_t)k * T] = (uint8_t)st;
            for (int t = 1; t < T; t++) {
                u = r.uniform();
                const double *row = A + 3 * st;
                st = u < row[0] ? 0 : (u < row[0] + row[1] ? 1 : 2);
                ss[(size_t)k * T + t] = (uint8_t)st;
            }
        }
        const size_t P = (size_t)X * Y * Z;
        float *dd = data + data_off[i];
        const double r2pi = std::sqrt(2 * M_PI), sd = std::sqrt(var);
        const int rxy = (int)std::ceil(5 * sigma), rz = (int)std::ceil(5 * sigma * zF);
        std::vector<double> img(P), gx, gy, gz;
        for (int t = 0; t < T; t++) {
            std::fill(img.begin(), img.end(), b0);
            for (int k = 0; k < Kb; k++) {
                if (ss[(size_t)k * T + t] != 0) continue;
                const double *ph = pp + 5 * k;
                const int x0 = std::max(0, (int)std::floor(ph[2]) - rxy), x1 = std::min(X - 1, (int)std::floor(ph[2]) + rxy + 1);
                const int y0 = std::max(0, (int)std::floor(ph[3]) - rxy), y1 = std::min(Y - 1, (int)std::floor(ph[3]) + rxy + 1);
                const int z0 = std::max(0, (int)std::floor(ph[4]) - rz), z1 = std::min(Z - 1, (int)std::floor(ph[4]) + rz + 1);
                const double amp = ph[0] / (ph[1] * r2pi), i2s = 1.0 / (2 * ph[1] * ph[1]);
                gx.assign(x1 - x0 + 1, 0);
                gy.assign(y1 - y0 + 1, 0);
                gz.assign(z1 - z0 + 1, 0);
                for (int x = x0; x <= x1; x++) gx[x - x0] = std::exp(-(x - ph[2]) * (x - ph[2]) * i2s);
                for (int y = y0; y <= y1; y++) gy[y - y0] = std::exp(-(y - ph[3]) * (y - ph[3]) * i2s);
                for (int z = z0; z <= z1; z++) {
                    const double dz = (z - ph[4]) / zF;
                    gz[z - z0] = amp * std::exp(-dz * dz * i2s);
                }
                for (int z = z0; z <= z1; z++)
                    for (int y = y0; y <= y1; y++) {
                        const double f = gz[z - z0] * gy[y - y0];
                        double *row = img.data() + ((size_t)z * Y + y) * X;
                        for (int x = x0; x <= x1; x++) row[x] += f * gx[x - x0];
                    }
            }
            float *dt = dd + (size_t)t * P;
            for (size_t p = 0; p < P; p += 2) {
                double u1 = r.uniform(), u2 = r.uniform();
                if (u1 < 1e-300) u1 = 1e-300;
                const double rad = std::sqrt(-2.0 * std::log(u1)), th = 2.0 * M_PI * u2;
                dt[p] = (float)(img[p] + sd * rad * std::cos(th));
                if (p + 1 < P) dt[p + 1] = (float)(img[p + 1] + sd * rad * std::sin(th));
            }
        }
    });
    return 0;
}

// Poisson draws for a given mean (sampler self-test).
int rfba_synth_poisson(double lam, uint64_t seed, int n, long *out) {
    Rng r(seed);
    for (int i = 0; i < n; i++) out[i] = poisson(r, lam);
    return 0;
}

}  // extern "C"
