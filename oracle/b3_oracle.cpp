// b3_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// FP64 host restatement of the 3B likelihood path, used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
// No product code path links or calls it.  Each function follows the cited
// reference lines in operation order (the gradient in forward mode, as the
// reference computes it; the GPU computes it in reverse mode).
//
// Pinning: every function is checked in tests/test_b3_oracle.py against
// finite differences, closed forms, brute-force enumeration of all state
// paths or the exact path posterior (see the test docstrings).  The
// reference's own 3B headers cannot be compiled here without stand-ins for
// MSVC-only names, and the reference ships no vectors for this path.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <thread>
#include <vector>

struct rfba3b_model {
    double A[9];
    double pi[3];
    double variance;
    double zfactor;
    int z_exact;
    int checkpoint;
    double mu_brightness, sigma_brightness;
    double mu_blur, sigma_blur;
};

namespace {

// ln, forward_algorithm.h:34-40
inline double ln(double x) { return x == 0 ? -1e100 : std::log(x); }
inline double sq(double x) { return x * x; }

// spot_shape_s + spot_shape_diff_position, storm.h:51-80.  With z_exact the
// z derivative carries the 1/zF^2 the reference omits (storm.h:78).
void spot_diff(const int *x, const double *phi, double zF, int z_exact, double &prob, double *deriv) {
    double d0 = x[0] - phi[2], d1 = x[1] - phi[3], d2 = x[2] - phi[4];
    d2 = d2 / zF;
    const double s = -(d0 * d0 + d1 * d1 + d2 * d2) / (2 * phi[1] * phi[1]);
    const double r_2_pi = std::sqrt(2 * M_PI);
    prob = std::exp(s) * phi[0] / (phi[1] * r_2_pi);
    const double f = std::exp(s) / (phi[1] * r_2_pi);
    deriv[0] = f * 1;
    deriv[1] = f * (-phi[0] * (1 + 2 * s) / phi[1]);
    deriv[2] = f * ((x[0] - phi[2]) * (phi[0] / sq(phi[1])));
    deriv[3] = f * ((x[1] - phi[3]) * (phi[0] / sq(phi[1])));
    double dz = (x[2] - phi[4]) * (phi[0] / sq(phi[1]));
    if (z_exact) dz = dz / (zF * zF);
    deriv[4] = f * dz;
}

// SpotWithBackground constructor, spot_with_background.hh:106-151.
void emissions(const double *px, const double *bg, int T, int P, const double *mu, const double *dmu, double var,
               double *on, double *off, double *diff) {
    for (int t = 0; t < T; t++) {
        double lo = 0, lf = 0, d[5] = {0, 0, 0, 0, 0};
        const double *pt = px + (size_t)t * P, *bt = bg + (size_t)t * P;
        for (int p = 0; p < P; p++) {
            const double e = pt[p] - (bt[p] + mu[p]);
            lf -= sq(pt[p] - bt[p]);
            lo -= sq(e);
            for (int i = 0; i < 5; i++) d[i] += dmu[5 * p + i] * e;
        }
        lo = lo / (2 * var) - P * std::log(2 * M_PI * var) / 2;
        lf = lf / (2 * var) - P * std::log(2 * M_PI * var) / 2;
        on[t] = lo;
        off[t] = lf;
        for (int i = 0; i < 5; i++) diff[5 * t + i] = d[i] / var;
    }
}

inline double B(int state, int t, const double *on, const double *off) { return state == 0 ? on[t] : off[t]; }

// forward_algorithm_hessian with compute_deriv, forward_algorithm.h:157-256.
void forward_deriv(const double *A, const double *pi, const double *on, const double *off, const double *diff, int T,
                   double &logL, double *grad) {
    std::vector<double> delta(3 * T), dd(15 * T);
    for (int i = 0; i < 3; i++) {
        delta[i] = ln(pi[i]) + B(i, 0, on, off);
        for (int n = 0; n < 5; n++) dd[5 * i + n] = (i == 0) ? diff[n] : 0.0;
    }
    for (int t = 1; t < T; t++)
        for (int j = 0; j < 3; j++) {
            double Z = -HUGE_VAL;
            for (int i = 0; i < 3; i++) Z = std::max(Z, delta[3 * (t - 1) + i] + ln(A[3 * i + j]));
            double sum = 0;
            for (int i = 0; i < 3; i++) sum += std::exp(delta[3 * (t - 1) + i] + ln(A[3 * i + j]) - Z);
            delta[3 * t + j] = B(j, t, on, off) + Z + ln(sum);
            double top[5] = {0, 0, 0, 0, 0};
            for (int i = 0; i < 3; i++) {
                const double w = std::exp(delta[3 * (t - 1) + i] + ln(A[3 * i + j]) - Z);
                for (int n = 0; n < 5; n++) top[n] += dd[15 * (t - 1) + 5 * i + n] * w;
            }
            for (int n = 0; n < 5; n++) dd[15 * t + 5 * j + n] = ((j == 0) ? diff[5 * t + n] : 0.0) + top[n] / sum;
        }
    double Z = -HUGE_VAL;
    for (int i = 0; i < 3; i++) Z = std::max(Z, delta[3 * (T - 1) + i]);
    double sum = 0;
    for (int i = 0; i < 3; i++) sum += std::exp(delta[3 * (T - 1) + i] - Z);
    logL = Z + ln(sum);
    for (int n = 0; n < 5; n++) grad[n] = 0;
    for (int i = 0; i < 3; i++) {
        const double w = std::exp(delta[3 * (T - 1) + i] - logL);
        for (int n = 0; n < 5; n++) grad[n] += w * dd[15 * (T - 1) + 5 * i + n];
    }
}

// forward_algorithm_delta, forward_algorithm.h:344-382 (+ the final log
// probability of forward_algorithm :301-310).
double forward_delta(const double *A, const double *pi, const double *on, const double *off, int T, double *delta) {
    for (int i = 0; i < 3; i++) delta[i] = ln(pi[i]) + B(i, 0, on, off);
    for (int t = 1; t < T; t++)
        for (int j = 0; j < 3; j++) {
            double Z = -HUGE_VAL;
            for (int i = 0; i < 3; i++) Z = std::max(Z, delta[3 * (t - 1) + i] + ln(A[3 * i + j]));
            double sum = 0;
            for (int i = 0; i < 3; i++) sum += std::exp(delta[3 * (t - 1) + i] + ln(A[3 * i + j]) - Z);
            delta[3 * t + j] = B(j, t, on, off) + Z + ln(sum);
        }
    double Z = -HUGE_VAL;
    for (int i = 0; i < 3; i++) Z = std::max(Z, delta[3 * (T - 1) + i]);
    double sum = 0;
    for (int i = 0; i < 3; i++) sum += std::exp(delta[3 * (T - 1) + i] - Z);
    return Z + ln(sum);
}

// select_random_element (forward_algorithm.h:502-513): the first i whose
// running sum reaches choice = u * scale, else the last element.
int select_random_element(const double *v, int n, double scale, double u) {
    double total = 0, choice = u * scale;
    for (int i = 0; i < n; i++) {
        total += v[i];
        if (choice <= total) return i;
    }
    return n - 1;
}

// sample_unscaled_log (forward_algorithm.h:520-532)
int sample_unscaled_log(const double *a_in, double u) {
    double a[3] = {a_in[0], a_in[1], a_in[2]};
    const double hi = std::max(a[0], std::max(a[1], a[2]));
    double sum = 0;
    for (int i = 0; i < 3; i++) {
        a[i] = std::exp(a[i] - hi);
        sum += a[i];
    }
    return select_random_element(a, 3, sum, u);
}

// backward_sampling (forward_algorithm.h:541-566): the k-th draw of the
// generator is u[k]; the last frame is drawn first.
void backward_sampling(const double *A, const double *delta, int T, const double *u, int *states) {
    double lA[9];
    for (int i = 0; i < 9; i++) lA[i] = ln(A[i]);
    int k = 0;
    states[T - 1] = sample_unscaled_log(delta + 3 * (T - 1), u[k++]);
    for (int t = T - 2; t >= 0; t--) {
        double rp[3];
        for (int j = 0; j < 3; j++) rp[j] = delta[3 * t + j] + lA[3 * j + states[t + 1]];
        states[t] = sample_unscaled_log(rp, u[k++]);
    }
}

// diff_log_log_normal, storm.h:434-437
inline double diff_log_log_normal(double x, double mu, double sigma) { return -(1 + (ln(x) - mu) / sq(sigma)) / x; }

void spot_tables(const int *xyz, int P, const double *phi, const rfba3b_model &m, std::vector<double> &mu,
                 std::vector<double> &dmu) {
    mu.resize(P);
    dmu.resize(5 * (size_t)P);
    for (int p = 0; p < P; p++) spot_diff(xyz + 3 * p, phi, m.zfactor, m.z_exact, mu[p], &dmu[5 * (size_t)p]);
}

template <typename F> void parallel_items(int n, int threads, F f) {
    threads = std::max(1, std::min(threads, n));
    std::vector<std::thread> th;
    for (int i = 0; i < threads; i++)
        th.emplace_back([=] {
            for (int j = i; j < n; j += threads) f(j);
        });
    for (auto &t : th) t.join();
}

// hess_log_log_normal, storm.h:446-449
inline double hess_log_log_normal(double x, double mu, double sigma) {
    return (1 + (ln(x) - mu - 1) / sq(sigma)) / sq(x);
}

// spot_shape_hess_position, storm.h:86-150: value, derivative and Hessian
// (row-major [5][5]) of the spot at one pixel.  z_exact = 0 keeps the
// reference's z terms, which use x[2] - phi[4] without the 1/zF^2 of the
// exact derivative (as the derivative does, storm.h:78), and takes the entry
// [4][3] as 0: the reference assigns that entry to itself (storm.h:143), so
// its value there is whatever the uninitialised matrix held.  z_exact = 1
// gives the exact second derivatives (symmetric): with s = -(dx^2 + dy^2 +
// dz^2/zF^2)/(2 sigma^2) and w = (dx, dy, dz/zF^2), d mu/d phi_i =
// I e^s w_i/(sigma^3 sqrt(2 pi)) and d2 mu/d phi_i d phi_j =
// I e^s (w_i w_j - sigma^2 c_i delta_ij)/(sigma^5 sqrt(2 pi)), c = (1, 1, 1/zF^2).
void spot_hess(const int *x, const double *phi, double zF, int z_exact, double &prob, double *deriv, double *hess) {
    double d0 = x[0] - phi[2], d1 = x[1] - phi[3], d2 = x[2] - phi[4];
    const double d2s = d2 / zF;
    const double s = -(d0 * d0 + d1 * d1 + d2s * d2s) / (2 * phi[1] * phi[1]);
    const double r_2_pi = std::sqrt(2 * M_PI);
    const double es = std::exp(s);
    const double I = phi[0], sg = phi[1];
    prob = es * I / (sg * r_2_pi);
    const double w[3] = {d0, d1, z_exact ? d2 / (zF * zF) : d2};
    const double c[3] = {1.0, 1.0, z_exact ? 1.0 / (zF * zF) : 1.0};
    const double f = es / (sg * r_2_pi);
    deriv[0] = f * 1;
    deriv[1] = f * (-I * (1 + 2 * s) / sg);
    for (int i = 0; i < 3; i++) deriv[2 + i] = f * (w[i] * (I / sq(sg)));
    double *h = hess;
    h[0] = 0;
    h[1] = h[5] = -es * (1 + 2 * s) / (sg * sg * r_2_pi);
    for (int i = 0; i < 3; i++) h[2 + i] = h[5 * (2 + i)] = es * w[i] / (std::pow(sg, 3) * r_2_pi);
    h[6] = 2 * I * es * (1 + 5 * s + 2 * s * s) / (std::pow(sg, 3) * r_2_pi);
    for (int i = 0; i < 3; i++)
        h[7 + i] = h[5 * (2 + i) + 1] = -I * es * (3 + 2 * s) * w[i] / (std::pow(sg, 4) * r_2_pi);
    for (int i = 0; i < 3; i++)
        for (int j = i; j < 3; j++)
            h[5 * (2 + i) + 2 + j] = h[5 * (2 + j) + 2 + i] =
                I * es * (w[i] * w[j] - (i == j ? sq(sg) * c[i] : 0.0)) / (r_2_pi * std::pow(sg, 5));
    if (!z_exact) h[5 * 4 + 3] = 0.0;
}

// The per-frame parts of hess_handler (multispot5.cc:839-855) for one
// background: hess_t = sum_p e H_p - d_p d_p^T and deriv_t = sum_p e d_p with
// e = data - (bg + mu), in pixel order.  hp [T][25], dp [T][5].
void hess_parts(const std::vector<double> &mu, const std::vector<double> &dmu, const std::vector<double> &hmu,
                const double *data, const double *bg, int T, int P, double *hp, double *dp) {
    for (int t = 0; t < T; t++) {
        double H[25] = {0}, D[5] = {0};
        for (int p = 0; p < P; p++) {
            const double e = data[(size_t)t * P + p] - (bg[(size_t)t * P + p] + mu[p]);
            const double *d = &dmu[5 * (size_t)p], *h = &hmu[25 * (size_t)p];
            for (int i = 0; i < 5; i++)
                for (int j = 0; j < 5; j++) H[5 * i + j] += e * h[5 * i + j] - d[i] * d[j];
            for (int i = 0; i < 5; i++) D[i] += e * d[i];
        }
        std::copy(H, H + 25, hp + 25 * (size_t)t);
        std::copy(D, D + 5, dp + 5 * (size_t)t);
    }
}

}  // namespace

extern "C" {

int oracle_b3_spot(const int *xyz, int P, const double *phi, double zF, int z_exact, double *mu, double *dmu) {
    for (int p = 0; p < P; p++) spot_diff(xyz + 3 * p, phi, zF, z_exact, mu[p], dmu + 5 * (size_t)p);
    return 0;
}

int oracle_b3_emissions(const double *px, const double *bg, int T, int P, const double *mu, const double *dmu,
                        double var, double *on, double *off, double *diff) {
    emissions(px, bg, T, P, mu, dmu, var, on, off, diff);
    return 0;
}

int oracle_b3_forward_deriv(const double *A, const double *pi, const double *on, const double *off,
                            const double *diff, int T, double *logL, double *grad) {
    forward_deriv(A, pi, on, off, diff, T, *logL, grad);
    return 0;
}

double oracle_b3_forward_delta(const double *A, const double *pi, const double *on, const double *off, int T,
                               double *delta) {
    return forward_delta(A, pi, on, off, T, delta);
}

double oracle_b3_diff_log_log_normal(double x, double mu, double sigma) { return diff_log_log_normal(x, mu, sigma); }

int oracle_b3_backward_sampling(const double *A, const double *delta, int T, const double *u, int *states) {
    backward_sampling(A, delta, T, u, states);
    return 0;
}

// logL and d logL/dphi of one spot over one background (data, bg [T][P]).
int oracle_b3_logl_grad(const int *xyz, int P, int T, const double *data, const double *bg, const double *phi,
                        const rfba3b_model *m, double *logL, double *grad) {
    std::vector<double> mu, dmu, on(T), off(T), diff(5 * (size_t)T);
    spot_tables(xyz, P, phi, *m, mu, dmu);
    emissions(data, bg, T, P, mu.data(), dmu.data(), m->variance, on.data(), off.data(), diff.data());
    forward_deriv(m->A, m->pi, on.data(), off.data(), diff.data(), T, *logL, grad);
    return 0;
}

// SpotNegProbabilityDiffWithSampledBackground::operator(), multispot5.cc:616-669
// (samples summed in index order).
int oracle_b3_neg_grad(const int *xyz, int P, int T, const double *data, const double *bg, int S, const double *phi,
                       const rfba3b_model *m, double *g, double *logL_s) {
    if (phi[0] <= 0 || phi[1] <= 0) {
        for (int n = 0; n < 5; n++) g[n] = std::numeric_limits<double>::quiet_NaN();
        return 0;
    }
    double sum[5] = {0, 0, 0, 0, 0};
    for (int s = 0; s < S; s++) {
        double L, gr[5];
        oracle_b3_logl_grad(xyz, P, T, data, bg + (size_t)s * T * P, phi, m, &L, gr);
        if (logL_s) logL_s[s] = L;
        for (int n = 0; n < 5; n++) sum[n] += gr[n];
    }
    double d[5];
    for (int n = 0; n < 5; n++) d[n] = sum[n] / S;
    const double prior[5] = {diff_log_log_normal(phi[0], m->mu_brightness, m->sigma_brightness),
                             diff_log_log_normal(phi[1], m->mu_blur, m->sigma_blur), 0, 0, 0};
    for (int n = 0; n < 5; n++) g[n] = -(d[n] + prior[n]);
    return 0;
}

// Batched model evaluation of one block: item (q, k, s) is spot k at
// params[q][k] over b0 + the sampled on-intensities of the other spots.
int oracle_b3_model_eval(const int *dims, int T, const double *data, double b0, const rfba3b_model *m, int K,
                         const double *params, int Q, const uint8_t *states, int S, double *logL, double *grad,
                         int threads) {
    const int X = dims[0], Y = dims[1], Z = dims[2];
    const int P = X * Y * Z;
    std::vector<int> xyz(3 * (size_t)P);
    for (int p = 0; p < P; p++) {
        xyz[3 * p] = p % X;
        xyz[3 * p + 1] = (p / X) % Y;
        xyz[3 * p + 2] = p / (X * Y);
    }
    parallel_items(Q * S, threads, [&](int qs) {
        const int q = qs / S, s = qs % S;
        std::vector<std::vector<double>> mus(K);
        std::vector<double> dmu;
        for (int j = 0; j < K; j++) spot_tables(xyz.data(), P, params + (size_t)(q * K + j) * 5, *m, mus[j], dmu);
        std::vector<double> bg((size_t)T * P);
        for (int k = 0; k < K; k++) {
            for (int t = 0; t < T; t++)
                for (int p = 0; p < P; p++) {
                    double v = b0;
                    for (int j = 0; j < K; j++)
                        if (j != k && states[((size_t)s * K + j) * T + t] == 0) v += mus[j][p];
                    bg[(size_t)t * P + p] = v;
                }
            double L, gr[5];
            oracle_b3_logl_grad(xyz.data(), P, T, data, bg.data(), params + (size_t)(q * K + k) * 5, m, &L, gr);
            const size_t item = ((size_t)q * K + k) * S + s;
            logL[item] = L;
            if (grad)
                for (int n = 0; n < 5; n++) grad[item * 5 + n] = gr[n];
        }
    });
    return 0;
}

int oracle_b3_spot_hess(const int *xyz, int P, const double *phi, double zF, int z_exact, double *mu, double *dmu,
                        double *hess) {
    for (int p = 0; p < P; p++)
        spot_hess(xyz + 3 * p, phi, zF, z_exact, mu[p], dmu + 5 * (size_t)p, hess + 25 * (size_t)p);
    return 0;
}

double oracle_b3_hess_log_log_normal(double x, double mu, double sigma) { return hess_log_log_normal(x, mu, sigma); }

// sampled_background_spot_hessian_ffbs (multispot5.cc:902-982) with the state
// paths given: paths [S][n][T] (0 = on) with weights [S][n].  Per background s
// and path, H = sum over on frames of hess_t and D = sum of deriv_t
// (hess_handler, :870-887); sum_hess = sum w (H + D D^T), sum_deriv = sum w D;
// then sum_hess / (S var) - (sum_deriv / (S var))(sum_deriv / (S var))^T plus
// the log-normal priors' second derivatives on [0][0] and [1][1]
// (:960-975).  The reference's paths come from backward sampling, each with
// weight 1 / bs_iterations; the exact posterior's weights give the limit.
int oracle_b3_hessian_paths(const int *xyz, int P, int T, const double *data, const double *bg, int S,
                            const double *phi, const rfba3b_model *m, int n, const uint8_t *paths,
                            const double *weights, double *hess_out, double *deriv_out) {
    std::vector<double> mu(P), dmu(5 * (size_t)P), hmu(25 * (size_t)P);
    for (int p = 0; p < P; p++)
        spot_hess(xyz + 3 * p, phi, m->zfactor, m->z_exact, mu[p], &dmu[5 * (size_t)p], &hmu[25 * (size_t)p]);
    std::vector<double> hp(25 * (size_t)T), dp(5 * (size_t)T);
    double sh[25] = {0}, sd[5] = {0};
    for (int s = 0; s < S; s++) {
        hess_parts(mu, dmu, hmu, data, bg + (size_t)s * T * P, T, P, hp.data(), dp.data());
        for (int k = 0; k < n; k++) {
            const uint8_t *path = paths + ((size_t)s * n + k) * T;
            const double w = weights[(size_t)s * n + k];
            double H[25] = {0}, D[5] = {0};
            for (int t = 0; t < T; t++)
                if (path[t] == 0) {
                    for (int i = 0; i < 25; i++) H[i] += hp[25 * (size_t)t + i];
                    for (int i = 0; i < 5; i++) D[i] += dp[5 * (size_t)t + i];
                }
            for (int i = 0; i < 5; i++)
                for (int j = 0; j < 5; j++) sh[5 * i + j] += w * (H[5 * i + j] + D[i] * D[j]);
            for (int i = 0; i < 5; i++) sd[i] += w * D[i];
        }
    }
    const double norm = (double)S * m->variance;
    for (int i = 0; i < 25; i++) sh[i] /= norm;
    for (int i = 0; i < 5; i++) sd[i] /= norm;
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) sh[5 * i + j] -= sd[i] * sd[j];
    sh[0] += hess_log_log_normal(phi[0], m->mu_brightness, m->sigma_brightness);
    sh[6] += hess_log_log_normal(phi[1], m->mu_blur, m->sigma_blur);
    std::copy(sh, sh + 25, hess_out);
    if (deriv_out) {
        sd[0] += diff_log_log_normal(phi[0], m->mu_brightness, m->sigma_brightness);
        sd[1] += diff_log_log_normal(phi[1], m->mu_blur, m->sigma_blur);
        std::copy(sd, sd + 5, deriv_out);
    }
    return 0;
}

// The whole of sampled_background_spot_hessian_ffbs: per background s the
// forward messages of the spot (SpotWithBackground with its values only,
// forward_algorithm_delta), then bs backward samplings, draw k of iteration
// it being uniforms[(s bs + it) T + k] (the last frame first).  states
// [S][bs][T] (may be NULL) receives the paths.
int oracle_b3_hessian_ffbs(const int *xyz, int P, int T, const double *data, const double *bg, int S,
                           const double *phi, const rfba3b_model *m, int bs, const double *uniforms, double *hess_out,
                           double *deriv_out, uint8_t *states) {
    std::vector<double> mu, dmu, on(T), off(T), diff(5 * (size_t)T), delta(3 * (size_t)T);
    spot_tables(xyz, P, phi, *m, mu, dmu);
    std::vector<uint8_t> paths((size_t)S * bs * T);
    std::vector<int> st(T);
    for (int s = 0; s < S; s++) {
        emissions(data, bg + (size_t)s * T * P, T, P, mu.data(), dmu.data(), m->variance, on.data(), off.data(),
                  diff.data());
        forward_delta(m->A, m->pi, on.data(), off.data(), T, delta.data());
        for (int it = 0; it < bs; it++) {
            backward_sampling(m->A, delta.data(), T, uniforms + ((size_t)s * bs + it) * T, st.data());
            for (int t = 0; t < T; t++) paths[((size_t)s * bs + it) * T + t] = (uint8_t)st[t];
        }
    }
    if (states) std::copy(paths.begin(), paths.end(), states);
    std::vector<double> w((size_t)S * bs, 1.0 / bs);
    return oracle_b3_hessian_paths(xyz, P, T, data, bg, S, phi, m, bs, paths.data(), w.data(), hess_out, deriv_out);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Gibbs sampling of the blink states.

// getBall (multispot5.cc:1256-1269): the offsets of a ball, z outermost,
// x innermost; the test compares Ref3D::mag_squared (unsigned,
// image3d.h:78-82) converted to float with radius * radius in float.
static std::vector<std::array<int, 3>> get_ball(float radius) {
    std::vector<std::array<int, 3>> pts;
    const int r = (int)std::ceil(radius + 1);
    for (int z = -r; z <= r; z++)
        for (int y = -r; y <= r; y++)
            for (int x = -r; x <= r; x++) {
                const unsigned int m2 = unsigned(x * x) + unsigned(y * y) + unsigned(z * z);
                if (m2 <= radius * radius) pts.push_back({x, y, z});
            }
    return pts;
}

// get_spot_pixels (multispot5.cc:1001-1020) with ir_rounded3D (:989-995):
// the ball of radius spot[1] * 6 + 1 around the rounded position, looked up
// in the pixel list (sorted by Ref3D::operator<, image3d.h:50-53) with
// lower_bound.
static void get_spot_pixels(const int *xyz, int P, const double *phi, std::vector<int> &out) {
    auto less = [](const std::array<int, 3> &a, const std::array<int, 3> &b) {
        return a[2] < b[2] || (a[2] == b[2] && a[1] < b[1]) || (a[2] == b[2] && a[1] == b[1] && a[0] < b[0]);
    };
    std::vector<std::array<int, 3>> pixels(P);
    for (int p = 0; p < P; p++) pixels[p] = {xyz[3 * p], xyz[3 * p + 1], xyz[3 * p + 2]};
    const std::vector<std::array<int, 3>> pix = get_ball((float)(phi[1] * 6 + 1));
    std::array<int, 3> offset;
    for (int j = 0; j < 3; j++) {
        const double v = phi[2 + j];
        offset[j] = static_cast<int>(v > 0.0 ? v + 0.5 : v - 0.5);
    }
    out.clear();
    for (const auto &b : pix) {
        const std::array<int, 3> q = {b[0] + offset[0], b[1] + offset[1], b[2] + offset[2]};
        const int pos = (int)(std::lower_bound(pixels.begin(), pixels.end(), q, less) - pixels.begin());
        if (pos != P && pixels[pos] == q) out.push_back(pos);
    }
}

// remove_spot / add_spot over a mask (sampled_multispot.h:113-131).
static void masked_spot(std::vector<double> &intens, int P, int T, const double *spot, const uint8_t *smp,
                        const int *mask, int nm, double sign) {
    for (int t = 0; t < T; t++)
        if (smp[t] == 0)
            for (int p = 0; p < nm; p++) {
                double &v = intens[(size_t)t * P + mask[p]];
                v = sign < 0 ? v - spot[mask[p]] : v + spot[mask[p]];
            }
}

// SpotWithBackgroundMasked (spot_with_background.hh:106-151 over the mask):
// per frame, log P(frame | on) and log P(frame | off).
static void masked_emissions(const double *px, const std::vector<double> &bg, int P, int T, const double *spot,
                             const int *mask, int nm, double var, double *on, double *off) {
    for (int t = 0; t < T; t++) {
        double lo = 0, lf = 0;
        for (int p = 0; p < nm; p++) {
            const size_t i = (size_t)t * P + mask[p];
            const double e = px[i] - (bg[i] + spot[mask[p]]);
            lf -= sq(px[i] - bg[i]);
            lo -= sq(e);
        }
        on[t] = lo / (2 * var) - nm * std::log(2 * M_PI * var) / 2;
        off[t] = lf / (2 * var) - nm * std::log(2 * M_PI * var) / 2;
    }
}

extern "C" {

int oracle_b3_spot_pixels(const int *xyz, int P, const double *phi, int K, int *offsets, int *idx, size_t cap) {
    std::vector<int> out;
    size_t tot = 0;
    offsets[0] = 0;
    for (int k = 0; k < K; k++) {
        get_spot_pixels(xyz, P, phi + 5 * (size_t)k, out);
        for (int v : out) {
            if (idx && tot < cap) idx[tot] = v;
            tot++;
        }
        offsets[k + 1] = (int)tot;
    }
    return (idx && tot > cap) ? -1 : 0;
}

// GibbsSampler2::next (sampled_multispot.h:681-709): `iters` sweeps over the
// spots in index order.  data, intens [T][P]; spot [K][P]; masks as offsets
// [K+1] + indices; sample [K][T] (0 on, 1 off, 2 bleached).  Draw k of the
// generator is u[k]; *used returns the number of draws consumed.
int oracle_b3_gibbs_next(int P, int T, int K, const double *data, const double *spot, const int *moff,
                         const int *midx, const double *A, const double *pi, double var, int iters, uint8_t *sample,
                         double *intens, const double *u, size_t *used) {
    std::vector<double> cur(intens, intens + (size_t)T * P);
    std::vector<double> on(T), off(T), delta(3 * (size_t)T);
    std::vector<int> st(T);
    size_t k_u = 0;
    for (int j = 0; j < iters; j++)
        for (int k = 0; k < K; k++) {
            const double *sp = spot + (size_t)k * P;
            const int *mask = midx + moff[k];
            const int nm = moff[k + 1] - moff[k];
            uint8_t *smp = sample + (size_t)k * T;
            masked_spot(cur, P, T, sp, smp, mask, nm, -1.0);
            masked_emissions(data, cur, P, T, sp, mask, nm, var, on.data(), off.data());
            forward_delta(A, pi, on.data(), off.data(), T, delta.data());
            backward_sampling(A, delta.data(), T, u + k_u, st.data());
            k_u += T;
            for (int t = 0; t < T; t++) smp[t] = (uint8_t)st[t];
            masked_spot(cur, P, T, sp, smp, mask, nm, +1.0);
        }
    std::copy(cur.begin(), cur.end(), intens);
    if (used) *used = k_u;
    return 0;
}

}  // extern "C"
