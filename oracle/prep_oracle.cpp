// prep_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Host restatement of RFBA's image normalisation, used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker; no
// product code path links or calls it.
//
//   preprocess_image          RFBA Main/storm_imagery_no_load.cc:44-79
//   convolveGaussian_fir      libcvd 0.8 (third-party, not in the reference
//                             tree: only its header is vendored).  Restated
//                             from the header's template convolveGaussian
//                             (include/cvd/convolution.h:400-499), which is
//                             the same FIR: a normalised Gaussian of half-width
//                             ceil(3 sigma), edges repeated, rows first.
//   mean_and_variance         storm_imagery_no_load.cc:87-110
//   load_and_normalize_images storm_imagery.cc:358-416
//
// Pinning (tests/test_prep_oracle.py): the blur against a literal Python
// restatement of the header template's sliding-window code (bit for bit, in
// float32), against scipy.ndimage.gaussian_filter (same truncation, nearest
// edges) within float rounding, and the high-pass on constant and linear
// images.  Parity unpinned: the FP32 rounding order of the absent libcvd
// source file, which is assumed to be the header template's.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct Kern {
    std::vector<float> k;  // k[i - 1] for tap distance i
    double factor;
    int ks;
};

// convolution.h:405-412
Kern make_kernel(double sigma, double sigmas) {
    Kern K;
    K.ks = (int)std::ceil(sigmas * sigma);
    K.k.resize(K.ks);
    float ksum = 0.0f;
    for (int i = 1; i <= K.ks; i++) ksum += (K.k[i - 1] = (float)std::exp(-i * i / (2 * sigma * sigma)));
    for (int i = 0; i < K.ks; i++) K.k[i] /= (2 * ksum + 1);
    K.factor = 1.0 / (2 * ksum + 1);
    return K;
}

// One w x h float image blurred: every row, then every column.  Row pixel j
// starts from (float)(x * factor) in double at the row's first and last
// ksize pixels and x * (float)factor in the middle (convolveMiddle,
// convolution.h:376-393); columns start from (float)(x * factor) in double
// (assign_multiple).  Then (near + far) * kernel[k] is added for each k,
// with indices past an edge replaced by the edge.
void blur(const float *in, int w, int h, const Kern &K, float *out) {
    const int ks = K.ks;
    const float ff = (float)K.factor;
    std::vector<float> rows((size_t)w * h);
    for (int y = 0; y < h; y++) {
        const float *r = in + (size_t)y * w;
        for (int j = 0; j < w; j++) {
            float s = (j < ks || j >= w - ks) ? (float)(r[j] * K.factor) : r[j] * ff;
            for (int k = 0; k < ks; k++) {
                const int a = std::max(j - k - 1, 0), b = std::min(j + k + 1, w - 1);
                s += (r[a] + r[b]) * K.k[k];
            }
            rows[(size_t)y * w + j] = s;
        }
    }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float s = (float)(rows[(size_t)y * w + x] * K.factor);
            for (int k = 0; k < ks; k++) {
                const int a = std::max(y - k - 1, 0), b = std::min(y + k + 1, h - 1);
                s += (rows[(size_t)a * w + x] + rows[(size_t)b * w + x]) * K.k[k];
            }
            out[(size_t)y * w + x] = s;
        }
}

// preprocess_image of one stack: per z-slice, slice - blur on the interior,
// 0 on the one-pixel border (the zero-initialised f, :60-64); or a copy.
void highpass(const float *in, int X, int Y, int Z, const Kern &K, int skip, float *out) {
    const size_t A = (size_t)X * Y;
    if (skip) {
        memcpy(out, in, A * Z * sizeof(float));
        return;
    }
    std::vector<float> wide(A);
    for (int z = 0; z < Z; z++) {
        const float *s = in + z * A;
        float *o = out + z * A;
        blur(s, X, Y, K, wide.data());
        for (size_t i = 0; i < A; i++) o[i] = 0.0f;
        for (int r = 1; r < Y - 1; r++)
            for (int c = 1; c < X - 1; c++) o[(size_t)r * X + c] = s[(size_t)r * X + c] - wide[(size_t)r * X + c];
    }
}

// mean_and_variance: double sums in image, z, y, x order of x and of the
// float square of x; the pair is returned as floats.
void mean_var(const float *v, size_t n, double *mv) {
    double sum = 0, sum2 = 0;
    for (size_t i = 0; i < n; i++) {
        sum += v[i];
        sum2 += v[i] * v[i];  // float square (utility.h:48)
    }
    sum /= (double)n;
    sum2 /= (double)n;
    mv[0] = (float)sum;
    mv[1] = (float)(sum2 - sum * sum);
}

}  // namespace

extern "C" {

int oracle_prep_kernel(double sigma, double sigmas, float *k, int cap, int *ks, double *factor) {
    const Kern K = make_kernel(sigma, sigmas);
    *ks = K.ks;
    *factor = K.factor;
    if (cap < K.ks) return -1;
    std::copy(K.k.begin(), K.k.end(), k);
    return 0;
}

int oracle_prep_blur(const float *in, int w, int h, double sigma, double sigmas, float *out) {
    blur(in, w, h, make_kernel(sigma, sigmas), out);
    return 0;
}

// load_and_normalize_images over a series already converted to float
// ([T][Z][Y][X]): preprocess every stack (stacks in parallel on `threads`
// threads), mean_and_variance, then x / sqrt(variance) or, with fixed
// scaling, (x - sub) / div; stats = {mean, var} before and after.
int oracle_prep_normalize(const float *in, int X, int Y, int Z, int T, int skip, double lpf, int fixed, double sub,
                          double div, float *out, double *stats, int threads) {
    const Kern K = skip ? Kern{{}, 1.0, 0} : make_kernel(lpf, 3.0);
    const size_t S = (size_t)X * Y * Z;
    threads = std::max(1, std::min(threads, T));
    std::vector<std::thread> th;
    for (int i = 0; i < threads; i++)
        th.emplace_back([=, &K] {
            for (int t = i; t < T; t += threads) highpass(in + t * S, X, Y, Z, K, skip, out + t * S);
        });
    for (auto &t : th) t.join();
    double mv[2];
    mean_var(out, S * T, mv);
    stats[0] = mv[0];
    stats[1] = mv[1];
    const double s = fixed ? div : std::sqrt(mv[1]);
    const double b = fixed ? sub : 0.0;
    for (size_t i = 0; i < S * T; i++) out[i] = fixed ? (float)((out[i] - b) / s) : (float)(out[i] / s);
    mean_var(out, S * T, mv);
    stats[2] = mv[0];
    stats[3] = mv[1];
    return 0;
}

}  // extern "C"
