// prep.hip -- librfba_prep.so: RFBA's image loading and normalisation on
// gfx950.  Host TIFF reading (img_load3D, storm_imagery.cc:185-254) and the
// device form of load_and_normalize_images (storm_imagery.cc:358-416):
// per z-slice high-pass (preprocess_image, storm_imagery_no_load.cc:44-79)
// with libcvd's Gaussian FIR, the series' mean and variance
// (mean_and_variance, :87-110) and the scaling.
//
// High-pass kernel: one workgroup per 64 x 32 output tile of one slice.  The
// tile's input window, clamped at the slice's edges (libcvd's edge rule:
// rows and columns past the edge repeat the edge), is staged in LDS; the
// horizontal pass writes the blurred rows of the window to LDS and the
// vertical pass forms the outputs.  The float arithmetic is that of
// convolveGaussian (libcvd convolution.h:400-499): the centre term first,
// then (left + right) * kernel[k] added for k = 0 .. ksize - 1, no fused
// multiply-add, with the centre weight applied in double and rounded to
// float at the row ends and in the vertical pass, and in float in the
// middle of a row (convolveMiddle, :376-393).  Per-tile sums of the outputs
// and their float squares (FP64) go to a partial array that one workgroup
// reduces in a fixed order.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/rfba_prep.h"

namespace {

thread_local char g_err[512] = "";
std::mutex g_mu;
int g_ready = 0;

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define PHIP(expr)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return fail(e_ == hipErrorOutOfMemory ? -2 : -3, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

int dev_init() {
    if (g_ready) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(-4, "no HIP device");
    int dev = 0;
    PHIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    PHIP(hipGetDeviceProperties(&prop, dev));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return fail(-4, "device %s is not gfx950", prop.gcnArchName);
    g_ready = 1;
    return 0;
}

// ---------------------------------------------------------------------------
// TIFF

struct File {
    FILE *f = nullptr;
    size_t size = 0;
    bool be = false;
    ~File() {
        if (f) fclose(f);
    }
    bool read_at(size_t off, void *dst, size_t n) {
        if (off > size || n > size - off) return false;
        if (fseeko(f, (off_t)off, SEEK_SET) != 0) return false;
        return fread(dst, 1, n, f) == n;
    }
    uint16_t u16(const uint8_t *p) const { return be ? (uint16_t)(p[0] << 8 | p[1]) : (uint16_t)(p[1] << 8 | p[0]); }
    uint32_t u32(const uint8_t *p) const {
        return be ? ((uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3])
                  : ((uint32_t)p[3] << 24 | (uint32_t)p[2] << 16 | (uint32_t)p[1] << 8 | p[0]);
    }
};

struct Page {
    uint32_t width = 0, height = 0, bits = 1, compression = 1, spp = 1, planar = 1, format = 1;
    uint32_t rows_per_strip = 0xFFFFFFFFu;
    std::vector<uint32_t> offsets, counts;
};

int open_file(const char *path, File &F) {
    if (!path) return fail(-1, "path is NULL");
    F.f = fopen(path, "rb");
    if (!F.f) return fail(-3, "cannot open %s", path);
    if (fseeko(F.f, 0, SEEK_END) != 0) return fail(-3, "cannot seek %s", path);
    F.size = (size_t)ftello(F.f);
    uint8_t h[8];
    if (!F.read_at(0, h, 8)) return fail(-3, "%s: too short for a TIFF header", path);
    if (h[0] == 'I' && h[1] == 'I') F.be = false;
    else if (h[0] == 'M' && h[1] == 'M') F.be = true;
    else return fail(-3, "%s: not a TIFF file", path);
    if (F.u16(h + 2) != 42) return fail(-3, "%s: not a classic TIFF (BigTIFF is not supported)", path);
    return 0;
}

// The values of one IFD entry (SHORT or LONG) into v.
int entry_values(File &F, const uint8_t *e, std::vector<uint32_t> &v, const char *path) {
    const uint16_t type = F.u16(e + 2);
    const uint32_t count = F.u32(e + 4);
    const size_t tsz = type == 3 ? 2 : type == 4 ? 4 : 0;
    if (!tsz) return fail(-3, "%s: tag %u has type %u (SHORT or LONG expected)", path, F.u16(e), type);
    if (count > (1u << 26)) return fail(-3, "%s: tag %u has %u values", path, F.u16(e), count);
    std::vector<uint8_t> buf(tsz * count);
    if (tsz * count <= 4) {
        memcpy(buf.data(), e + 8, tsz * count);
    } else if (!F.read_at(F.u32(e + 8), buf.data(), buf.size())) {
        return fail(-3, "%s: tag %u points outside the file", path, F.u16(e));
    }
    v.resize(count);
    for (uint32_t i = 0; i < count; i++) v[i] = tsz == 2 ? F.u16(&buf[2 * i]) : F.u32(&buf[4 * i]);
    return 0;
}

// Every page's directory (the IFD chain, TIFFReadDirectory's order).
int read_pages(File &F, const char *path, std::vector<Page> &pages) {
    uint8_t h[8];
    F.read_at(0, h, 8);
    uint32_t off = F.u32(h + 4);
    while (off) {
        if (pages.size() > (1u << 20)) return fail(-3, "%s: too many pages", path);
        uint8_t nb[2];
        if (!F.read_at(off, nb, 2)) return fail(-3, "%s: directory outside the file", path);
        const uint16_t n = F.u16(nb);
        std::vector<uint8_t> ent(12 * (size_t)n + 4);
        if (!F.read_at(off + 2, ent.data(), ent.size())) return fail(-3, "%s: truncated directory", path);
        Page p;
        for (uint16_t i = 0; i < n; i++) {
            const uint8_t *e = &ent[12 * (size_t)i];
            const uint16_t tag = F.u16(e);
            std::vector<uint32_t> v;
            switch (tag) {
            case 256: case 257: case 258: case 259: case 273: case 277: case 278: case 279: case 284: case 339:
                if (int rc = entry_values(F, e, v, path)) return rc;
                if (v.empty()) return fail(-3, "%s: tag %u has no value", path, tag);
                break;
            default:
                continue;
            }
            switch (tag) {
            case 256: p.width = v[0]; break;
            case 257: p.height = v[0]; break;
            case 258: p.bits = v[0]; break;
            case 259: p.compression = v[0]; break;
            case 273: p.offsets = v; break;
            case 277: p.spp = v[0]; break;
            case 278: p.rows_per_strip = v[0]; break;
            case 279: p.counts = v; break;
            case 284: p.planar = v[0]; break;
            case 339: p.format = v[0]; break;
            }
        }
        if (p.compression != 1) return fail(-3, "%s: page %zu is compressed (%u)", path, pages.size(), p.compression);
        if (p.spp != 1) return fail(-3, "%s: page %zu has %u samples per pixel", path, pages.size(), p.spp);
        if (p.bits != 8 && p.bits != 16) return fail(-3, "%s: page %zu has %u bits per sample", path, pages.size(), p.bits);
        if (p.format == 3) return fail(-3, "%s: page %zu holds floating-point samples", path, pages.size());
        if (!p.width || !p.height) return fail(-3, "%s: page %zu has no size", path, pages.size());
        // page sizes and byte counts stay far inside int and size_t
        if (p.width > (1u << 20) || p.height > (1u << 20))
            return fail(-3, "%s: page %zu is %u x %u pixels (at most 2^20 per side)", path, pages.size(), p.width,
                        p.height);
        if (p.offsets.empty() || p.offsets.size() != p.counts.size())
            return fail(-3, "%s: page %zu has inconsistent strips", path, pages.size());
        pages.push_back(std::move(p));
        off = F.u32(&ent[12 * (size_t)n]);
    }
    if (pages.empty()) return fail(-3, "%s: no pages", path);
    for (size_t i = 1; i < pages.size(); i++)
        if (pages[i].width != pages[0].width || pages[i].height != pages[0].height || pages[i].bits != pages[0].bits)
            return fail(-3, "%s: page %zu differs in size or depth from page 0", path, i);
    return 0;
}

// ---------------------------------------------------------------------------
// kernels

constexpr int kTX = 64, kTY = 32;  // output tile (x, y)
constexpr int kThreads = 256;

struct Fir {
    float kern[64];    // kernel[k], k < ksize
    float factor_f;    // centre weight as float (middle of a row)
    int ksize;
    double factor_d;   // centre weight as double
};

__device__ __forceinline__ float to_f(uint8_t v) { return (float)v; }
__device__ __forceinline__ float to_f(uint16_t v) { return (float)v; }
__device__ __forceinline__ float to_f(float v) { return v; }

// KS > 0: the half-width fixed at compile time, each thread computing 8
// consecutive outputs of a row (then of a column) from a window of
// 8 + 2 KS values held in registers; KS = 0: any half-width, one output at a
// time from LDS.  Both evaluate the same terms in the same order.
constexpr int kRun = 8;
constexpr int kHS = kTX + 1;  // row stride of the blurred rows in LDS

template <typename Tin, int KS>
__global__ __launch_bounds__(kThreads) void prep_highpass(const Tin *__restrict__ raw, int X, int Y, int ntx,
                                                           int nty, long long ntiles, Fir F, int skip,
                                                           float *__restrict__ out, double2 *__restrict__ partial) {
    extern __shared__ float lds[];
    const int ks = KS > 0 ? KS : F.ksize;
    const int WX = kTX + 2 * ks, WY = kTY + 2 * ks;
    const int WS = WX | 1;        // odd row strides: lanes on consecutive rows hit distinct banks
    float *win = lds;             // [WY][WS] input window
    float *hrow = lds + WY * WS;  // [WY][kHS] horizontally blurred rows
    const int tid = threadIdx.x;
    // Workgroup b runs on XCD b % 8: each XCD takes a contiguous range of
    // tiles, so that neighbouring tiles' overlapping windows meet in its L2
    // (the grid is padded to a multiple of 8).
    const long long tile = (long long)(blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    if (tile >= ntiles) return;
    const int tx = (int)(tile % ntx), ty = (int)((tile / ntx) % nty);
    const long long slice = tile / ((long long)ntx * nty);
    const int x0 = tx * kTX, y0 = ty * kTY;
    const Tin *src = raw + slice * (long long)X * Y;

    // Interior uint16 tiles (no edge clamping, rows 8-byte aligned): each row
    // of the window from aligned 4-value loads starting at x0 - ks - 1
    // (x0 is a multiple of kTX, a multiple of 4, and ks = 15), 4 values
    // per load instead of one.
    constexpr bool kVec = std::is_same<Tin, uint16_t>::value && KS == 15;
    const int xa = x0 - ks - 1;
    if (kVec && xa >= 0 && xa + WX + 2 <= X && y0 - ks >= 0 && y0 + kTY + ks <= Y && (X & 3) == 0 &&
        ((size_t)src & 7) == 0) {
        constexpr int NV4 = (KS > 0 ? kTX + 2 * KS + 2 : 4) / 4;  // 4-value loads per row (96 values)
        for (int e = tid; e < NV4 * WY; e += kThreads) {
            const int r = e / NV4, v = e - r * NV4;
            const ushort4 q = *(const ushort4 *)(src + (size_t)(y0 - ks + r) * X + xa + 4 * v);
            const int c = 4 * v - 1;  // window column of q.x
            float *wr = win + r * WS;
            if (c >= 0) wr[c] = to_f((Tin)q.x);
            wr[c + 1] = to_f((Tin)q.y);
            wr[c + 2] = to_f((Tin)q.z);
            if (c + 3 < WX) wr[c + 3] = to_f((Tin)q.w);
        }
    } else {
        for (int e = tid; e < WX * WY; e += kThreads) {
            const int r = e / WX, c = e - r * WX;
            const int gy = min(max(y0 - ks + r, 0), Y - 1), gx = min(max(x0 - ks + c, 0), X - 1);
            win[r * WS + c] = to_f(src[(size_t)gy * X + gx]);
        }
    }
    __syncthreads();
    if (!skip) {
        if (KS > 0) {
            // rows: (row r, columns c0 .. c0 + 7) per item, lanes on consecutive rows
            for (int e = tid; e < WY * (kTX / kRun); e += kThreads) {
                const int c0 = (e / WY) * kRun, r = e - (e / WY) * WY;
                float w[kRun + 2 * (KS > 0 ? KS : 1)];
                const float *in = win + r * WS + c0;
#pragma unroll
                for (int i = 0; i < kRun + 2 * KS; i++) w[i] = in[i];
#pragma unroll
                for (int i = 0; i < kRun; i++) {
                    const int j = x0 + c0 + i;
                    float h;
                    if (j < KS || j >= X - KS) h = (float)((double)w[KS + i] * F.factor_d);
                    else h = w[KS + i] * F.factor_f;
#pragma unroll
                    for (int k = 0; k < KS; k++) h += (w[KS + i - k - 1] + w[KS + i + k + 1]) * F.kern[k];
                    hrow[r * kHS + c0 + i] = h;
                }
            }
        } else {
            for (int e = tid; e < WY * kTX; e += kThreads) {
                const int r = e / kTX, c = e - r * kTX;
                const int j = x0 + c;
                const float *in = win + r * WS + c + ks;
                float h;
                if (j < ks || j >= X - ks) h = (float)((double)in[0] * F.factor_d);
                else h = in[0] * F.factor_f;
                for (int k = 0; k < ks; k++) h += (in[-k - 1] + in[k + 1]) * F.kern[k];
                hrow[r * kHS + c] = h;
            }
        }
        __syncthreads();
    }
    double s1 = 0.0, s2 = 0.0;
    float *dst = out + slice * (long long)X * Y;
    auto emit = [&](int r, int c, float v) {
        const int gy = y0 + r, gx = x0 + c;
        if (gy >= Y || gx >= X) return;
        const float x = win[(r + ks) * WS + c + ks];
        float f;
        if (skip) f = x;
        else if (gy >= 1 && gy <= Y - 2 && gx >= 1 && gx <= X - 2) f = x - v;
        else f = 0.0f;
        dst[(size_t)gy * X + gx] = f;
        s1 += (double)f;
        s2 += (double)(f * f);
    };
    if (KS > 0 && !skip) {
        // columns: (column c, rows r0 .. r0 + 7) per item
        for (int e = tid; e < kTX * (kTY / kRun); e += kThreads) {
            const int c = e % kTX, r0 = (e / kTX) * kRun;
            float w[kRun + 2 * (KS > 0 ? KS : 1)];
            const float *hc = hrow + r0 * kHS + c;
#pragma unroll
            for (int i = 0; i < kRun + 2 * KS; i++) w[i] = hc[i * kHS];
#pragma unroll
            for (int i = 0; i < kRun; i++) {
                float v = (float)((double)w[KS + i] * F.factor_d);
#pragma unroll
                for (int k = 0; k < KS; k++) v += (w[KS + i - k - 1] + w[KS + i + k + 1]) * F.kern[k];
                emit(r0 + i, c, v);
            }
        }
    } else {
        for (int e = tid; e < kTY * kTX; e += kThreads) {
            const int r = e / kTX, c = e - r * kTX;
            float v = 0.0f;
            if (!skip) {
                const float *hc = hrow + (r + ks) * kHS + c;
                v = (float)((double)hc[0] * F.factor_d);
                for (int k = 0; k < ks; k++) v += (hc[-(k + 1) * kHS] + hc[(k + 1) * kHS]) * F.kern[k];
            }
            emit(r, c, v);
        }
    }
    __shared__ double red[2][kThreads / 64];
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
    }
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = s1;
        red[1][tid >> 6] = s2;
    }
    __syncthreads();
    if (tid == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kThreads / 64; w++) {
            a += red[0][w];
            b += red[1][w];
        }
        partial[tile] = make_double2(a, b);
    }
}

// (x - sub) / div (fixed scaling) or x / div (div = sqrt of the variance),
// in double, rounded to float; per-block sums of the results and their
// float squares.  Block b covers elements [b * chunk, (b + 1) * chunk), chunk
// a multiple of 4 (16-byte accesses; the tail past a multiple of 4 singly).
__global__ __launch_bounds__(kThreads) void prep_scale(float *__restrict__ v, long long n, long long chunk,
                                                       double sub, double div, double2 *__restrict__ partial) {
    const long long b0 = (long long)blockIdx.x * chunk, b1 = max(b0, min(n, b0 + chunk));  // empty past n
    double s1 = 0.0, s2 = 0.0;
    const long long n4 = b0 + ((b1 - b0) & ~3LL);
    for (long long i = b0 + 4 * (long long)threadIdx.x; i < n4; i += 4 * kThreads) {
        float4 q = *(const float4 *)(v + i);
        float r[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            r[j] = (float)(((double)r[j] - sub) / div);
            s1 += (double)r[j];
            s2 += (double)(r[j] * r[j]);
        }
        *(float4 *)(v + i) = make_float4(r[0], r[1], r[2], r[3]);
    }
    for (long long i = n4 + threadIdx.x; i < b1; i += kThreads) {
        const float f = (float)(((double)v[i] - sub) / div);
        v[i] = f;
        s1 += (double)f;
        s2 += (double)(f * f);
    }
    __shared__ double red[2][kThreads / 64];
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = s1;
        red[1][threadIdx.x >> 6] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kThreads / 64; w++) {
            a += red[0][w];
            b += red[1][w];
        }
        partial[blockIdx.x] = make_double2(a, b);
    }
}

// Sums of n partials in a fixed order: block b of gridDim.x sums the
// contiguous range [b * chunk, (b + 1) * chunk) into out[b] (a double2 per
// block), so that a second launch with one block finishes the total.
__global__ __launch_bounds__(1024) void prep_reduce(const double2 *__restrict__ partial, long long n, long long chunk,
                                                    double2 *__restrict__ out) {
    const long long b0 = (long long)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
    double s1 = 0.0, s2 = 0.0;
    for (long long i = b0 + threadIdx.x; i < b1; i += 1024) {
        s1 += partial[i].x;
        s2 += partial[i].y;
    }
    __shared__ double red[2][16];
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = s1;
        red[1][threadIdx.x >> 6] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < 16; w++) {
            a += red[0][w];
            b += red[1][w];
        }
        out[blockIdx.x] = make_double2(a, b);
    }
}

// ---------------------------------------------------------------------------
// host helpers

int gaussian(double sigma, double sigmas, Fir &F) {
    if (!(sigma > 0) || !(sigmas > 0) || !std::isfinite(sigma * sigmas)) return fail(-1, "sigma must be positive");
    const int ksize = (int)std::ceil(sigmas * sigma);
    if (ksize > 64) return fail(-1, "filter half-width %d exceeds 64 (sigma %g)", ksize, sigma);
    float ksum = 0.0f;
    for (int i = 1; i <= ksize; i++) ksum += (F.kern[i - 1] = (float)std::exp(-i * i / (2 * sigma * sigma)));
    for (int i = 0; i < ksize; i++) F.kern[i] /= (2 * ksum + 1);
    F.factor_d = 1.0 / (2 * ksum + 1);
    F.factor_f = (float)F.factor_d;
    F.ksize = ksize;
    return 0;
}

// mean_and_variance's result from the two sums: (float) mean, (float) variance
void stats_of(const double *s, double area, double *out) {
    const double m = s[0] / area, q = s[1] / area;
    out[0] = (float)m;
    out[1] = (float)(q - m * m);
}

constexpr int kRed = 256;  // blocks of the first reduction level

struct Scratch {
    double2 *partial = nullptr;
    size_t cap = 0;
    double2 *sums = nullptr;  // [kRed] first level, then [kRed] the total
} g_scr;

// Two launches: kRed blocks over n partials, then one block over kRed.
void reduce2(long long n, hipStream_t st) {
    const long long chunk = (n + kRed - 1) / kRed;
    hipLaunchKernelGGL(prep_reduce, dim3(kRed), dim3(1024), 0, st, (const double2 *)g_scr.partial, n, chunk, g_scr.sums);
    hipLaunchKernelGGL(prep_reduce, dim3(1), dim3(1024), 0, st, (const double2 *)g_scr.sums, (long long)kRed,
                       (long long)kRed, g_scr.sums + kRed);
}

int scratch(size_t n) {
    if (!g_scr.sums) PHIP(hipMalloc(&g_scr.sums, 2 * kRed * sizeof(double2)));
    if (n > g_scr.cap) {
        if (g_scr.partial) PHIP(hipFree(g_scr.partial));
        g_scr.partial = nullptr;
        g_scr.cap = 0;
        PHIP(hipMalloc(&g_scr.partial, n * sizeof(double2)));
        g_scr.cap = n;
    }
    return 0;
}

int normalize(const void *d_raw, int type, const int *dims, const rfba_prep_conf *conf, float *d_out, double *stats,
              hipStream_t st, float *kernel_ms) {
    if (!d_raw || !dims || !conf || !d_out) return fail(-1, "NULL argument");
    const int X = dims[0], Y = dims[1], Z = dims[2], T = dims[3];
    if (X <= 0 || Y <= 0 || Z <= 0 || T <= 0) return fail(-1, "dims must be positive");
    if (type != RFBA_PREP_U8 && type != RFBA_PREP_U16 && type != RFBA_PREP_F32) return fail(-1, "unknown type %d", type);
    if (conf->fixed_scaling && !conf->skip)
        return fail(-1, "fixed scaling needs preprocess.skip (load_and_normalize_images exits otherwise)");
    if (conf->fixed_scaling && !(conf->divide != 0 && std::isfinite(conf->divide) && std::isfinite(conf->subtract)))
        return fail(-1, "fixed scaling needs a finite nonzero divide and a finite subtract");
    Fir F;
    memset(&F, 0, sizeof F);
    if (!conf->skip) {
        if (int rc = gaussian(conf->lpf, 3.0, F)) return rc;
        if (X < 2 * F.ksize || Y < 2 * F.ksize + 1)
            return fail(-1, "slice %d x %d is smaller than the blur needs (%d x %d for sigma %g)", X, Y, 2 * F.ksize,
                        2 * F.ksize + 1, (double)conf->lpf);
    }
    if (int rc = dev_init()) return rc;
    const long long ntx = (X + kTX - 1) / kTX, nty = (Y + kTY - 1) / kTY;
    const long long ntiles = ntx * nty * (long long)Z * T;
    if (ntiles > 0x7ffffff0LL) return fail(-1, "series too large for one launch");
    const long long n = (long long)X * Y * Z * T;
    const int sblocks = 4096;
    const long long chunk = ((n + sblocks - 1) / sblocks + 3) & ~3LL;  // a multiple of 4 elements
    if (int rc = scratch(std::max<long long>(ntiles, sblocks))) return rc;
    const size_t lds = 4 * (size_t)(kTY + 2 * F.ksize) * (((size_t)(kTX + 2 * F.ksize) | 1) + kHS);
    // the default blur (sigma 5: half-width 15) runs the register-window form
    const bool k15 = !conf->skip && F.ksize == 15;
    typedef void (*HpFn)(const void *, int, int, int, int, long long, Fir, int, float *, double2 *);
    HpFn fn;
    if (type == RFBA_PREP_U8) fn = k15 ? (HpFn)prep_highpass<uint8_t, 15> : (HpFn)prep_highpass<uint8_t, 0>;
    else if (type == RFBA_PREP_U16) fn = k15 ? (HpFn)prep_highpass<uint16_t, 15> : (HpFn)prep_highpass<uint16_t, 0>;
    else fn = k15 ? (HpFn)prep_highpass<float, 15> : (HpFn)prep_highpass<float, 0>;
    if (lds > 64 * 1024) PHIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));

    hipEvent_t e0, e1;
    PHIP(hipEventCreate(&e0));
    PHIP(hipEventCreate(&e1));
    PHIP(hipEventRecord(e0, st));
    const int skip = conf->skip ? 1 : 0;
    hipLaunchKernelGGL(fn, dim3((unsigned)((ntiles + 7) / 8 * 8)), dim3(kThreads), lds, st, d_raw, X, Y, (int)ntx,
                       (int)nty, ntiles, F, skip, d_out, g_scr.partial);
    PHIP(hipGetLastError());
    reduce2(ntiles, st);
    PHIP(hipGetLastError());
    double sums[2], before[2], after[2];
    PHIP(hipMemcpyAsync(sums, g_scr.sums + kRed, sizeof sums, hipMemcpyDeviceToHost, st));
    PHIP(hipStreamSynchronize(st));
    const double area = (double)n;
    stats_of(sums, area, before);
    // load_and_normalize_images: (x - subtract) / divide, or x / sqrt(variance)
    const double sub = conf->fixed_scaling ? conf->subtract : 0.0;
    const double div = conf->fixed_scaling ? conf->divide : std::sqrt(before[1]);
    hipLaunchKernelGGL(prep_scale, dim3(sblocks), dim3(kThreads), 0, st, d_out, n, chunk, sub, div, g_scr.partial);
    PHIP(hipGetLastError());
    reduce2(sblocks, st);
    PHIP(hipGetLastError());
    PHIP(hipMemcpyAsync(sums, g_scr.sums + kRed, sizeof sums, hipMemcpyDeviceToHost, st));
    PHIP(hipEventRecord(e1, st));
    PHIP(hipEventSynchronize(e1));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    stats_of(sums, area, after);
    if (stats) {
        stats[0] = before[0];
        stats[1] = before[1];
        stats[2] = after[0];
        stats[3] = after[1];
    }
    if (kernel_ms) *kernel_ms = ms;
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" const char *rfba_prep_last_error(void) { return g_err; }

extern "C" int rfba_tiff_info_read(const char *path, rfba_tiff_info *info) {
    if (!info) return fail(-1, "info is NULL");
    File F;
    if (int rc = open_file(path, F)) return rc;
    std::vector<Page> pages;
    if (int rc = read_pages(F, path, pages)) return rc;
    info->width = (int)pages[0].width;
    info->height = (int)pages[0].height;
    info->pages = (int)pages.size();
    info->bits = (int)pages[0].bits;
    return 0;
}

extern "C" int rfba_tiff_load(const char *path, void *out, size_t cap, rfba_tiff_info *info) {
    File F;
    if (int rc = open_file(path, F)) return rc;
    std::vector<Page> pages;
    if (int rc = read_pages(F, path, pages)) return rc;
    const size_t bpp = pages[0].bits / 8;
    const size_t page_bytes = (size_t)pages[0].width * pages[0].height * bpp;
    const size_t total = page_bytes * pages.size();
    if (info) {
        info->width = (int)pages[0].width;
        info->height = (int)pages[0].height;
        info->pages = (int)pages.size();
        info->bits = (int)pages[0].bits;
    }
    if (!out) return fail(-1, "output is NULL");
    if (total > cap) return fail(-1, "%s needs %zu bytes, capacity %zu", path, total, cap);
    uint8_t *dst = (uint8_t *)out;
    for (size_t p = 0; p < pages.size(); p++) {
        size_t got = 0;
        for (size_t s = 0; s < pages[p].offsets.size() && got < page_bytes; s++) {
            const size_t n = std::min<size_t>(pages[p].counts[s], page_bytes - got);
            if (!F.read_at(pages[p].offsets[s], dst + p * page_bytes + got, n))
                return fail(-3, "%s: page %zu strip %zu outside the file", path, p, s);
            got += n;
        }
        if (got != page_bytes) return fail(-3, "%s: page %zu holds %zu of %zu bytes", path, p, got, page_bytes);
    }
    if (bpp == 2 && F.be) {
        uint16_t *v = (uint16_t *)out;
        for (size_t i = 0; i < total / 2; i++) v[i] = (uint16_t)(v[i] << 8 | v[i] >> 8);
    }
    return 0;
}

extern "C" int rfba_tiff_write(const char *path, const void *px, int width, int height, int pages, int bits,
                               int big_endian, int rows_per_strip) {
    if (!path || !px || width <= 0 || height <= 0 || pages <= 0 || (bits != 8 && bits != 16))
        return fail(-1, "invalid TIFF write arguments");
    const int rps = rows_per_strip > 0 ? std::min(rows_per_strip, height) : height;
    const int nstrips = (height + rps - 1) / rps;
    const size_t bpp = bits / 8, row = (size_t)width * bpp, page_bytes = row * height;
    std::vector<uint8_t> out;
    auto put16 = [&](size_t at, uint32_t v) {
        if (big_endian) { out[at] = (uint8_t)(v >> 8); out[at + 1] = (uint8_t)v; }
        else { out[at] = (uint8_t)v; out[at + 1] = (uint8_t)(v >> 8); }
    };
    auto put32 = [&](size_t at, uint32_t v) {
        for (int i = 0; i < 4; i++) out[at + (big_endian ? 3 - i : i)] = (uint8_t)(v >> (8 * i));
    };
    out.resize(8);
    out[0] = out[1] = big_endian ? 'M' : 'I';
    put16(2, 42);
    size_t link = 4;  // where the next IFD offset goes
    const uint8_t *src = (const uint8_t *)px;
    for (int p = 0; p < pages; p++) {
        // pixel data first, then the strip tables, then the IFD
        const size_t data0 = out.size();
        out.resize(data0 + page_bytes);
        memcpy(&out[data0], src + p * page_bytes, page_bytes);
        if (bits == 16 && big_endian)
            for (size_t i = 0; i < page_bytes; i += 2) std::swap(out[data0 + i], out[data0 + i + 1]);
        if (out.size() & 1) out.push_back(0);
        const size_t offs = out.size();
        out.resize(offs + 8 * (size_t)nstrips);
        for (int s = 0; s < nstrips; s++) {
            const int r0 = s * rps, nr = std::min(rps, height - r0);
            put32(offs + 4 * s, (uint32_t)(data0 + (size_t)r0 * row));
            put32(offs + 4 * ((size_t)nstrips + s), (uint32_t)((size_t)nr * row));
        }
        const size_t ifd = out.size();
        const int n = 9;
        out.resize(ifd + 2 + 12 * n + 4, 0);
        put16(ifd, n);
        size_t e = ifd + 2;
        auto entry = [&](uint16_t tag, uint16_t type, uint32_t count, uint32_t value) {
            put16(e, tag);
            put16(e + 2, type);
            put32(e + 4, count);
            if (type == 3 && count == 1) put16(e + 8, value);
            else put32(e + 8, value);
            e += 12;
        };
        entry(256, 4, 1, (uint32_t)width);
        entry(257, 4, 1, (uint32_t)height);
        entry(258, 3, 1, (uint32_t)bits);
        entry(259, 3, 1, 1);
        entry(262, 3, 1, 1);
        entry(273, 4, (uint32_t)nstrips, nstrips == 1 ? (uint32_t)data0 : (uint32_t)offs);
        entry(277, 3, 1, 1);
        entry(278, 4, 1, (uint32_t)rps);
        entry(279, 4, (uint32_t)nstrips, nstrips == 1 ? (uint32_t)page_bytes : (uint32_t)(offs + 4 * nstrips));
        put32(link, (uint32_t)ifd);
        link = ifd + 2 + 12 * n;
        if (out.size() > 0xFFFFFFF0u) return fail(-1, "TIFF larger than 4 GiB");
    }
    FILE *f = fopen(path, "wb");
    if (!f) return fail(-3, "cannot create %s", path);
    const size_t w = fwrite(out.data(), 1, out.size(), f);
    fclose(f);
    return w == out.size() ? 0 : fail(-3, "short write to %s", path);
}

extern "C" int rfba_prep_gaussian_kernel(double sigma, double sigmas, float *kernel, int cap, int *ksize,
                                         double *factor) {
    Fir F;
    if (int rc = gaussian(sigma, sigmas, F)) return rc;
    if (ksize) *ksize = F.ksize;
    if (factor) *factor = F.factor_d;
    if (kernel) {
        if (cap < F.ksize) return fail(-1, "kernel needs %d entries", F.ksize);
        memcpy(kernel, F.kern, F.ksize * sizeof(float));
    }
    return 0;
}

extern "C" int rfba_prep_normalize_device(const void *d_raw, int type, const int *dims, const rfba_prep_conf *conf,
                                          float *d_out, double *stats, void *stream, float *kernel_ms) {
    std::lock_guard<std::mutex> lk(g_mu);
    return normalize(d_raw, type, dims, conf, d_out, stats, (hipStream_t)stream, kernel_ms);
}

extern "C" int rfba_prep_load_and_normalize(const char *const *paths, int n, const rfba_prep_conf *conf, float *d_out,
                                            size_t cap, int *dims, double *stats, void *stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!paths || n <= 0 || !conf || !d_out) return fail(-1, "NULL argument or no files");
    if (int rc = dev_init()) return rc;
    rfba_tiff_info info;
    if (int rc = rfba_tiff_info_read(paths[0], &info)) return rc;
    const size_t stack = (size_t)info.width * info.height * info.pages;
    const size_t bytes = stack * (info.bits / 8);
    if (stack * (size_t)n > cap) return fail(-1, "%d stacks of %zu voxels exceed the output (%zu floats)", n, stack, cap);
    hipStream_t st = (hipStream_t)stream;
    void *d_raw = nullptr;
    void *pin[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    int rc = 0;
    auto cleanup = [&]() {
        for (int i = 0; i < 2; i++) {
            if (done[i]) (void)hipEventDestroy(done[i]);
            if (pin[i]) (void)hipHostFree(pin[i]);
        }
        if (d_raw) (void)hipFree(d_raw);
    };
#define LTRY(expr)                                                                                         \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) {                                                                            \
            rc = fail(e_ == hipErrorOutOfMemory ? -2 : -3, "%s: %s", #expr, hipGetErrorString(e_));        \
            cleanup();                                                                                     \
            return rc;                                                                                     \
        }                                                                                                  \
    } while (0)
    LTRY(hipMalloc(&d_raw, bytes * (size_t)n));
    for (int i = 0; i < 2; i++) {
        LTRY(hipHostMalloc(&pin[i], bytes, hipHostMallocDefault));
        LTRY(hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
    }
    // file i is read into pinned buffer i % 2 while file i - 1 is copied
    for (int i = 0; i < n; i++) {
        const int b = i & 1;
        if (i >= 2) LTRY(hipEventSynchronize(done[b]));
        rfba_tiff_info fi;
        if ((rc = rfba_tiff_load(paths[i], pin[b], bytes, &fi)) != 0) {
            cleanup();
            return rc;
        }
        if (fi.width != info.width || fi.height != info.height || fi.pages != info.pages || fi.bits != info.bits) {
            cleanup();
            return fail(-1, "%s differs in size from %s: all images must be the same size", paths[i], paths[0]);
        }
        LTRY(hipMemcpyAsync((uint8_t *)d_raw + (size_t)i * bytes, pin[b], bytes, hipMemcpyHostToDevice, st));
        LTRY(hipEventRecord(done[b], st));
    }
#undef LTRY
    const int d4[4] = {info.width, info.height, info.pages, n};
    rc = normalize(d_raw, info.bits == 8 ? RFBA_PREP_U8 : RFBA_PREP_U16, d4, conf, d_out, stats, st, nullptr);
    if (rc == 0 && dims) memcpy(dims, d4, sizeof d4);
    if (rc == 0) {
        const hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = fail(-3, "hipStreamSynchronize: %s", hipGetErrorString(e));
    }
    cleanup();
    return rc;
}

extern "C" int rfba_prep_malloc(void **p, size_t bytes) {
    if (!p) return fail(-1, "NULL pointer");
    if (int rc = dev_init()) return rc;
    PHIP(hipMalloc(p, bytes));
    return 0;
}

extern "C" int rfba_prep_free(void *p) {
    PHIP(hipFree(p));
    return 0;
}

extern "C" int rfba_prep_memcpy(void *dst, const void *src, size_t bytes, int kind) {
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost
                            : kind == 3 ? hipMemcpyDeviceToDevice : hipMemcpyDefault;
    if (kind < 1 || kind > 3) return fail(-1, "unknown copy kind %d", kind);
    PHIP(hipMemcpy(dst, src, bytes, k));
    return 0;
}
