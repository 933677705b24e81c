// srrf_api.cpp -- host side of librfba_srrf.so: the C ABI declared in
// include/rfba_srrf.h, SRRF.ini parsing, sampling-set analysis, tile and
// LDS planning, device buffer caching and kernel launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "srrf_internal.h"

namespace rfba {
template <typename Tin, int TZ, bool RP2> __global__ void srrf_fused(LaunchParams P);
__global__ void srrf_reset(int *counter, int *err);
__global__ void copy_probe(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n);
__global__ void srrf_clamp_init(unsigned *maxbits);
__global__ void srrf_clamp(float *__restrict__ d, size_t n, unsigned *__restrict__ maxbits);
__global__ void srrf_gather_region(const float *__restrict__ sr, int X, int Y, int M, const int *__restrict__ px,
                                   long long n, float *__restrict__ out);
}  // namespace rfba

using namespace rfba;

// ---------------------------------------------------------------------------
// errors
static thread_local char g_err[512] = "";

static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(e_ == hipErrorOutOfMemory ? RFBA_ENOMEM : RFBA_EDEVICE, "%s: %s", \
                        #expr, hipGetErrorString(e_));                                    \
    } while (0)

static int hip_status(hipError_t e) {
    if (e == hipSuccess) return RFBA_OK;
    return fail(e == hipErrorOutOfMemory ? RFBA_ENOMEM : RFBA_EDEVICE, "%s", hipGetErrorString(e));
}

extern "C" const char *rfba_last_error(void) { return g_err; }

// ---------------------------------------------------------------------------
// SRRF.ini (GradientCalc.cpp:15-42).  Same section and key names, same
// defaults.  Parsing follows GetPrivateProfile*: case-insensitive section and
// key match, first occurrence wins, ';' starts a comment line, integers are
// parsed like atoi (leading sign and digits), missing keys take the default.
namespace {

std::string lower(std::string s) {
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}
std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

struct Ini {
    std::vector<std::pair<std::string, std::pair<std::string, std::string>>> kv;  // sec,(key,val)
    bool load(const char *path) {
        FILE *f = fopen(path, "rb");
        if (!f) return false;
        std::string sec;
        char line[4096];
        while (fgets(line, sizeof line, f)) {
            std::string l = trim(line);
            if (l.empty() || l[0] == ';') continue;
            if (l[0] == '[') {
                size_t e = l.find(']');
                sec = lower(trim(l.substr(1, e == std::string::npos ? std::string::npos : e - 1)));
                continue;
            }
            size_t eq = l.find('=');
            if (eq == std::string::npos) continue;
            std::string k = lower(trim(l.substr(0, eq)));
            std::string v = trim(l.substr(eq + 1));
            if (v.size() >= 2 && ((v.front() == '"' && v.back() == '"') || (v.front() == '\'' && v.back() == '\'')))
                v = v.substr(1, v.size() - 2);
            kv.push_back({sec, {k, v}});
        }
        fclose(f);
        return true;
    }
    const std::string *find(const char *sec, const char *key) const {
        std::string s = lower(sec), k = lower(key);
        for (auto &e : kv)
            if (e.first == s && e.second.first == k) return &e.second.second;
        return nullptr;
    }
    int get_int(const char *sec, const char *key, int def) const {
        const std::string *v = find(sec, key);
        return v ? atoi(v->c_str()) : def;
    }
    std::string get_str(const char *sec, const char *key, const char *def) const {
        const std::string *v = find(sec, key);
        return v ? *v : std::string(def);
    }
};

}  // namespace

extern "C" void rfba_srrf_default_conf(SRRFConf *c) {
    c->RingRadius = 0.5f;
    c->Magnification = 5;
    c->AxesinRing = 6;
    c->MainRange = 1;
    c->XRange = 0;
    c->YRange = 0;
    c->ZRange = 0;
    c->Method = 1;
    c->ITC = 1;
    c->TRACOrder = 2;
    c->RemoveCon = 0;
    c->Renormalize = 0;
    c->GradientSmooth = 0;
    c->IntensityW = 1;
    c->blockSizeX = 16;
    c->blockSizeY = 16;
    c->blockSizeZ = 16;
}

extern "C" int rfba_srrf_load_ini(const char *path, SRRFConf *c) {
    if (!c) return fail(RFBA_EINVAL, "conf is NULL");
    Ini ini;
    ini.load(path ? path : "./SRRF.ini");
    c->RingRadius = (float)atof(ini.get_str("SRRF", "RingRadius", "0.5").c_str());
    c->Magnification = ini.get_int("SRRF", "Magnification", 5);
    c->AxesinRing = ini.get_int("SRRF", "AxesinRing", 6);
    c->MainRange = ini.get_int("SRRF", "MainRange", 1);
    c->XRange = ini.get_int("SRRF", "XRange", 0);
    c->YRange = ini.get_int("SRRF", "YRange", 0);
    c->ZRange = ini.get_int("SRRF", "ZRange", 0);
    c->Method = ini.get_int("Temproal", "Method", 1);
    c->ITC = ini.get_int("Temproal", "Integrate_Temporal_Correlations", 1);
    c->TRACOrder = ini.get_int("Temproal", "TRAC_ORDER", 2);
    c->RemoveCon = ini.get_int("Radiality", "Remove_Positivity_Constraint", 0);
    c->Renormalize = ini.get_int("Radiality", "Renormalize", 0);
    c->GradientSmooth = ini.get_int("Radiality", "Gradient_Smoothing", 0);
    c->IntensityW = ini.get_int("Weighting", "Intensity_Weighting", 1);
    c->blockSizeX = ini.get_int("Split", "BlockSizeX", 16);
    c->blockSizeY = ini.get_int("Split", "BlockSizeY", 16);
    c->blockSizeZ = ini.get_int("Split", "BlockSizeZ", 16);
    return RFBA_OK;
}

// ---------------------------------------------------------------------------
// sampling sets

extern "C" int rfba_srrf_build_ring(int axes, float radius, float *xyz) {
    // buildRing, kernel.cu:23-33: float angle, float sin/cos.
    const int n = 2 * axes;
    if (axes <= 0 || n * n > RFBA_SRRF_MAX_SAMPLES)
        return fail(RFBA_EINVAL, "AxesinRing %d out of range", axes);
    if (!xyz) return fail(RFBA_EINVAL, "xyz is NULL");
    const float PI = 3.141592654f;
    const float step = (PI * 2.0f) / (float)n;
    for (int a1 = 0; a1 < n; a1++)
        for (int a2 = 0; a2 < n; a2++) {
            const int i = a1 * n + a2;
            xyz[3 * i + 0] = radius * std::sin(step * a1) * std::cos(step * a2);
            xyz[3 * i + 1] = radius * std::sin(step * a1) * std::sin(step * a2);
            xyz[3 * i + 2] = radius * std::cos(step * a1);
        }
    return n * n;
}

extern "C" int rfba_srrf_build_dirs(int kind, float radius, float *xyz) {
    if (!xyz) return fail(RFBA_EINVAL, "xyz is NULL");
    if (kind != 6 && kind != 26) return fail(RFBA_EINVAL, "direction kind must be 6 or 26");
    struct P3 { float x, y, z; };
    std::vector<P3> v;
    for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dz = -1; dz <= 1; dz++) {
                const int nz = (dx != 0) + (dy != 0) + (dz != 0);
                if (nz == 0 || (kind == 6 && nz != 1)) continue;
                const double len = std::sqrt((double)(dx * dx + dy * dy + dz * dz));
                v.push_back({(float)(radius * dx / len), (float)(radius * dy / len), (float)(radius * dz / len)});
            }
    // Order: by |x| (largest first), then x, y, z.  The samples of one x
    // offset are consecutive (one group, runs of equal y), and the groups
    // x = -R and x = +R are neighbours: at R = 1/2 the kernel serves both
    // from one x pass (their sample positions differ by one SR column).
    std::stable_sort(v.begin(), v.end(), [](const P3 &a, const P3 &b) {
        if (std::fabs(a.x) != std::fabs(b.x)) return std::fabs(a.x) > std::fabs(b.x);
        if (a.x != b.x) return a.x < b.x;
        if (a.y != b.y) return a.y < b.y;
        return a.z < b.z;
    });
    for (size_t i = 0; i < v.size(); i++) {
        xyz[3 * i] = v[i].x;
        xyz[3 * i + 1] = v[i].y;
        xyz[3 * i + 2] = v[i].z;
    }
    return (int)v.size();
}

namespace {

int find_or_add(float *tab, int &n, float v) {
    for (int i = 0; i < n; i++)
        if (tab[i] == v) return i;  // +0 == -0: same coordinate
    if (n >= kMaxOffsets) return -1;
    tab[n] = v;
    return n++;
}

int analyse_samples(const float *xyz, int n, float divisor, float R, DevSamples &s) {
    if (n <= 0 || n > kMaxSamples) return fail(RFBA_EINVAL, "sample count %d out of range", n);
    if (!(divisor == divisor) || divisor == 0.0f) return fail(RFBA_EINVAL, "invalid divisor");
    memset(&s, 0, sizeof(s));
    s.n = n;
    s.divisor = divisor;
    s.R = R;
    int ex = 0;
    const float m = std::frexp(R, &ex);
    s.r_pow2 = (R > 0.0f && std::isnormal(R) && m == 0.5f) ? 1 : 0;
    s.invR = s.r_pow2 ? 1.0f / R : 0.0f;
    s.fast_ok = (R >= 0x1p-20f && R <= 0x1p20f) ? 1 : 0;
    s.oxmin = s.oymin = s.ozmin = 0.0f;
    s.oxmax = s.oymax = s.ozmax = 0.0f;
    int prev_ix = -1, prev_iy = -1;
    for (int i = 0; i < n; i++) {
        const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z))
            return fail(RFBA_EINVAL, "non-finite sample offset");
        const int ix = find_or_add(s.ox, s.nox, x);
        const int iy = find_or_add(s.oy, s.noy, y);
        const int iz = find_or_add(s.oz, s.noz, z);
        if (ix < 0 || iy < 0 || iz < 0) return fail(RFBA_EINVAL, "too many distinct offsets per axis");
        s.rx[i] = x;
        s.ry[i] = y;
        s.rz[i] = z;
        s.siz[i] = iz;
        s.oxmin = std::min(s.oxmin, x); s.oxmax = std::max(s.oxmax, x);
        s.oymin = std::min(s.oymin, y); s.oymax = std::max(s.oymax, y);
        s.ozmin = std::min(s.ozmin, z); s.ozmax = std::max(s.ozmax, z);
        if (ix == prev_ix && iy == prev_iy) {
            s.run_len[s.nruns - 1]++;
        } else {
            s.run_start[s.nruns] = i;
            s.run_len[s.nruns] = 1;
            s.run_iy[s.nruns] = iy;
            if (ix == prev_ix) {
                s.grp_nrun[s.ngroups - 1]++;
            } else {
                s.grp_ix[s.ngroups] = ix;
                s.grp_run0[s.ngroups] = s.nruns;
                s.grp_nrun[s.ngroups] = 1;
                s.ngroups++;
            }
            s.nruns++;
        }
        prev_ix = ix;
        prev_iy = iy;
    }
    // Merge consecutive groups a, b = a + 1 (x offsets multiples of 1/2, so
    // that (X + 1/2) + b == ((X + 1) + 1/2) + a exactly for every column X):
    // b's runs read a's planes one column over.  The runs stay in list order.
    int ng = 0;
    for (int g = 0; g < s.ngroups; g++) {
        const float a = s.ox[s.grp_ix[g]];
        bool merge = false;
        if (g + 1 < s.ngroups) {
            const float b = s.ox[s.grp_ix[g + 1]];
            merge = b - a == 1.0f && std::fabs(a) <= 0x1p20f && 2.0f * a == std::nearbyint(2.0f * a);
        }
        s.mgrp_ix[ng] = s.grp_ix[g];
        s.mgrp_run0[ng] = s.grp_run0[g];
        s.mgrp_nrun[ng] = s.grp_nrun[g];
        s.mgrp_ext[ng] = 0;
        for (int r = s.grp_run0[g]; r < s.grp_run0[g] + s.grp_nrun[g]; r++) s.run_shift[r] = 0;
        if (merge) {
            for (int r = s.grp_run0[g + 1]; r < s.grp_run0[g + 1] + s.grp_nrun[g + 1]; r++) s.run_shift[r] = 1;
            s.mgrp_nrun[ng] += s.grp_nrun[g + 1];
            s.mgrp_ext[ng] = 1;
            s.any_ext = 1;
            g++;
        }
        ng++;
    }
    s.nmgroups = ng;
    return RFBA_OK;
}

int samples_from(const SRRFConf *conf, const rfba_srrf_samples *smp, DevSamples &s) {
    if (smp) {
        if (!smp->xyz) return fail(RFBA_EINVAL, "samples->xyz is NULL");
        return analyse_samples(smp->xyz, smp->n, smp->divisor, conf->RingRadius, s);
    }
    const int nr = 2 * conf->AxesinRing;
    if (conf->AxesinRing <= 0 || nr * nr > kMaxSamples)
        return fail(RFBA_EINVAL, "AxesinRing %d out of range", conf->AxesinRing);
    std::vector<float> xyz(3 * nr * nr);
    rfba_srrf_build_ring(conf->AxesinRing, conf->RingRadius, xyz.data());
    return analyse_samples(xyz.data(), nr * nr, (float)nr, conf->RingRadius, s);
}

int check_conf(const SRRFConf *c) {
    if (!c) return fail(RFBA_EINVAL, "conf is NULL");
    if (c->Magnification <= 0 || c->Magnification > 64)
        return fail(RFBA_EINVAL, "Magnification %d out of range", c->Magnification);
    if (c->Method < 0 || c->Method > 3) return fail(RFBA_EINVAL, "unknown Method %d", c->Method);
    if (c->Method == 3 && (c->TRACOrder < 1 || c->TRACOrder > 4))
        return fail(RFBA_EINVAL, "TRACOrder %d out of range", c->TRACOrder);
    if (!(c->RingRadius == c->RingRadius)) return fail(RFBA_EINVAL, "RingRadius is NaN");
    if (c->GradientSmooth == 1 &&
        (c->MainRange < 0 || c->XRange < 0 || c->YRange < 0 || c->ZRange < 0 || c->MainRange > 8 ||
         c->XRange > 8 || c->YRange > 8 || c->ZRange > 8))
        return fail(RFBA_EINVAL, "gradient ranges out of range");
    return RFBA_OK;
}

// ---------------------------------------------------------------------------
// device context: cached buffers, one device, calls serialised.

struct Ctx {
    std::mutex mu;
    bool init = false;
    int dev = 0;
    int ncu = 0;
    DevSamples *d_smp = nullptr;
    int *d_ctl = nullptr;  // [0] counter, [1] err
    DevItem *d_items = nullptr;
    size_t items_cap = 0;
    DevTile *d_tiles = nullptr;
    size_t tiles_cap = 0;
    float *d_scratch = nullptr;
    size_t scratch_bytes = 0;
    void *d_in = nullptr;
    size_t in_bytes = 0;
    float *d_out = nullptr;
    size_t out_bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};
Ctx g_ctx;

int ctx_init() {
    if (g_ctx.init) return RFBA_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(RFBA_ENODEV, "no HIP device");
    HIP_TRY(hipGetDevice(&g_ctx.dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, g_ctx.dev));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(RFBA_ENODEV, "device %s is not gfx950", prop.gcnArchName);
    g_ctx.ncu = prop.multiProcessorCount;
    HIP_TRY(hipMalloc(&g_ctx.d_smp, sizeof(DevSamples)));
    HIP_TRY(hipMalloc(&g_ctx.d_ctl, 2 * sizeof(int)));
    HIP_TRY(hipEventCreate(&g_ctx.ev0));
    HIP_TRY(hipEventCreate(&g_ctx.ev1));
    g_ctx.init = true;
    return RFBA_OK;
}

template <typename T> int ensure(T *&p, size_t &cap, size_t bytes) {
    if (bytes <= cap && p) return RFBA_OK;
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    HIP_TRY(hipMalloc((void **)&p, bytes));
    cap = bytes;
    return RFBA_OK;
}

// ---------------------------------------------------------------------------
// planning

int first_tap_h(float c, int M) { return (int)std::floor(c / (float)M - 0.5f) - 1; }

struct Plan {
    LaunchGeom g;
    int merge = 0;  // merged sample groups (LaunchParams::merge)
    int TZ;
    size_t lds_bytes;
    int Tmax;
    std::vector<DevTile> tiles;             // queue order
    std::vector<std::pair<int, int>> item_tiles;  // per item [begin, count) in tiles_by_item
    std::vector<DevTile> tiles_by_item;
    std::vector<std::array<int, 3>> phase;  // per item: tile phase k of x, y, z (origins b0 - k + j w)
};

// Low-res taps spanned by the widest tile of one axis whose tiles of width tw
// start at b0 - k + j tw (j = 0, 1, ... while below b1), the kernel's own
// extent computation.
void axis_extent(int b0, int b1, int k, int tw, float omin, float omax, int M, int &need) {
    for (int X0 = b0 - k; X0 < b1; X0 += tw) {
        const int lo = first_tap_h(((float)X0 + 0.5f) + omin, M);
        const int hi = first_tap_h(((float)(X0 + tw - 1) + 0.5f) + omax, M) + 3;
        need = std::max(need, hi - lo + 1);
    }
}

// SR voxels [v0, v1) of [b0, b1) none of whose sample offsets along this axis
// leaves [1.5, dim - 1.5] low-res pixels (the kernel's float test; monotone
// in the coordinate).  The others are NaN on every frame and skip the sample
// passes.  Empty: v0 >= v1.
void axis_valid(int b0, int b1, const float *off, int noff, int M, int dim, int &v0, int &v1) {
    v0 = b1;
    v1 = b0;
    for (int X = b0; X < b1; X++) {
        const float Xc = (float)X + 0.5f;
        bool ok = true;
        for (int i = 0; i < noff && ok; i++) {
            const float x = (Xc + off[i]) / (float)M;
            ok = !(x < 1.5f || x > (float)dim - 1.5f);
        }
        if (ok) {
            v0 = std::min(v0, X);
            v1 = X + 1;
        }
    }
}

// Tile phases of one item.  Along each axis the tiles can start at any of tw
// phases; the cost of a tile-frame is modelled from the measured phases of
// the kernel: a tile with a voxel to sample pays the halo load, the centre
// intensity and the temporal accumulation (0.18), the gradient and the x
// passes (0.20) and the sample passes (0.62), which a wave (an 8 x 8 quadrant
// of the tile's columns) skips per pair of z slices when none of its voxels
// has samples in range.  A tile without one (idle) pays idle_cost: 0.18, or
// almost nothing where the kernel skips its frame body (integer input or no
// intensity weighting).  The costs are separable per axis, so
// each axis contributes its tile count n, its tiles with a valid voxel e and
// its active fraction a (quadrant halves for x and y, slice pairs for z).
// Only phases whose tiles span no more low-res taps than phase 0's are
// tried, so the LDS plan never grows; ties keep phase 0.
std::array<int, 3> tile_phases(const DevItem &it, const DevSamples &s, int TZ, double idle_cost) {
    struct Opt {
        double n, e, a;
        int k;
    };
    const int b0[3] = {it.bx0, it.by0, it.bz0}, b1[3] = {it.bx1, it.by1, it.bz1};
    const int tw[3] = {kTX, kTY, TZ}, gran[3] = {8, 8, 2}, dim[3] = {it.W, it.H, it.D};
    const float *off[3] = {s.ox, s.oy, s.oz};
    const int noff[3] = {s.nox, s.noy, s.noz};
    const float omin[3] = {s.oxmin, s.oymin, s.ozmin}, omax[3] = {s.oxmax, s.oymax, s.ozmax};
    std::vector<Opt> opts[3];
    for (int a = 0; a < 3; a++) {
        int v0, v1, need0 = 0;
        axis_valid(b0[a], b1[a], off[a], noff[a], it.M, dim[a], v0, v1);
        axis_extent(b0[a], b1[a], 0, tw[a], omin[a], omax[a], it.M, need0);
        for (int k = 0; k < tw[a]; k++) {
            int need = 0;
            axis_extent(b0[a], b1[a], k, tw[a], omin[a], omax[a], it.M, need);
            if (need > need0) continue;
            Opt o = {0.0, 0.0, 0.0, k};
            for (int X0 = b0[a] - k; X0 < b1[a]; X0 += tw[a]) {
                o.n += 1.0;
                const int lo = std::max(X0, v0), hi = std::min(X0 + tw[a], v1);
                if (lo < hi) {
                    o.e += 1.0;
                    o.a += (double)((hi - 1 - X0) / gran[a] - (lo - X0) / gran[a] + 1) / (double)(tw[a] / gran[a]);
                }
            }
            // phases with the same counts cost the same: keep the first
            bool seen = false;
            for (const auto &q : opts[a]) seen = seen || (q.n == o.n && q.e == o.e && q.a == o.a);
            if (!seen) opts[a].push_back(o);
        }
    }
    std::array<int, 3> best = {0, 0, 0};
    double bc = -1.0;
    for (const auto &x : opts[0])
        for (const auto &y : opts[1])
            for (const auto &z : opts[2]) {
                const double e = x.e * y.e * z.e;
                const double c = idle_cost * (x.n * y.n * z.n - e) + 0.38 * e + 0.62 * x.a * y.a * z.a;
                if (bc < 0.0 || c < bc) {
                    bc = c;
                    best = {x.k, y.k, z.k};
                }
            }
    return best;
}

size_t lds_layout(LaunchGeom &g, int TZ, const DevSamples &s, bool merge) {
    const int noz = s.noz;
    g.RU = g.NU + 2 * g.hx;
    g.RV = g.NV + 2 * g.hy;
    g.RW = g.NW + 2 * g.hz;
    // X stride of the x-interpolated planes: the y pass of a wave reads 8
    // consecutive x values from several consecutive v rows, so consecutive
    // rows are shifted by 8, 16 or 24 of the 32 banks (2 to 4 rows
    // disjoint); the least padding that does so.
    // A merged group's x pass writes kTX + 1 columns.
    const int xcols = kTX + (merge ? 1 : 0);
    g.TXP = xcols;
    for (int p = 0; p < 32; p++) {
        const int sh = (g.NW * (xcols + p)) % 32;
        if (sh == 8 || sh == 16 || sh == 24) {
            g.TXP = xcols + p;
            break;
        }
    }
    g.noz = noz;
    g.NWP = g.NW + (g.NW & 1);
    g.VS = (merge && g.NWP > 16 ? kTX + 1 : kTX) * g.NWP + 16;
    // Column layout when it costs no occupancy tier (workgroups per CU by
    // LDS: budgets of 53, 80, 160 KiB) against the row layout.
    auto tier = [](size_t words) {
        const size_t b = words * sizeof(float) + 64;
        return b <= 53 * 1024 ? 0 : (b <= 80 * 1024 ? 1 : 2);
    };
    const size_t rest = (size_t)3 * g.NU * g.NV * g.NW + 2 + std::max(3 * g.NW * kThreads, g.RU * g.RV * g.RW) +
                        xcols * 6 + TZ * noz * 8 + TZ + TZ * 6 + TZ * kThreads + s.nox + s.noy + s.noz +
                        4 * s.ngroups + 6 * s.nruns + 2 * s.n;
    const size_t px_row = (size_t)3 * g.NV * g.NW * g.TXP, px_col = (size_t)3 * g.NV * g.VS;
    g.pxv = tier(rest + px_col) <= tier(rest + px_row) ? 1 : 0;
    int off = 0;
    // The raw halo is read only before the sample passes (gradient and the
    // voxel-centre intensity); the y-interpolated columns reuse its space.
    g.off_g = off; off += 3 * g.NU * g.NV * g.NW;
    off += off & 1;  // 8-byte aligned planes (pairs of w in the column layout)
    g.off_px = off; off += g.pxv ? (int)px_col : (int)px_row;
    g.off_py = off;
    g.off_raw = off; off += std::max(3 * g.NW * kThreads, g.RU * g.RV * g.RW);
    g.off_xtab = off; off += xcols * 6;
    g.off_ztab = off; off += TZ * noz * 8;
    g.off_zc = off; off += TZ;
    g.off_ctab = off; off += TZ * 6;
    g.off_dsum = off; off += TZ * kThreads;
    g.off_stab = off;
    {
        int w = s.nox;
        g.st_oy = w; w += s.noy;
        g.st_oz = w; w += s.noz;
        g.st_grp = w; w += 4 * s.ngroups;
        g.st_run = w; w += 6 * s.nruns;
        g.st_smp = w; w += 2 * s.n;
        off += w;
    }
    g.off_total = off;
    return (size_t)off * sizeof(float);
}

int plan_geometry(const std::vector<DevItem> &items, const DevSamples &s, const SRRFConf &c, int dtype, int TZ,
                  Plan &pl) {
    LaunchGeom g;
    memset(&g, 0, sizeof g);
    if (c.GradientSmooth == 0) {
        g.hx = g.hy = g.hz = 1;
    } else {
        g.hx = std::max(c.MainRange, c.XRange);
        g.hy = std::max(c.MainRange, c.YRange);
        g.hz = std::max(c.MainRange, c.ZRange);
    }
    int NU = 4, NV = 4, NW = 4, Tmax = 1;
    pl.phase.resize(items.size());
    for (size_t i = 0; i < items.size(); i++) {
        const auto &it = items[i];
        const auto ph = tile_phases(it, s, TZ, (dtype == RFBA_U16 || c.IntensityW != 1) ? 0.005 : 0.18);
        pl.phase[i] = ph;
        axis_extent(it.bx0, it.bx1, ph[0], kTX, s.oxmin, s.oxmax, it.M, NU);
        axis_extent(it.by0, it.by1, ph[1], kTY, s.oymin, s.oymax, it.M, NV);
        axis_extent(it.bz0, it.bz1, ph[2], TZ, s.ozmin, s.ozmax, it.M, NW);
        Tmax = std::max(Tmax, it.T);
    }
    g.NU = NU; g.NV = NV; g.NW = NW;
    pl.lds_bytes = lds_layout(g, TZ, s, false) + 64;
    pl.merge = 0;
    if (s.any_ext) {
        // merged groups save x passes but widen the planes by a column: use
        // them only where that costs no workgroups per CU
        LaunchGeom gm = g;
        const size_t mb = lds_layout(gm, TZ, s, true) + 64;
        auto wg = [](size_t b) { return b <= 53 * 1024 ? 3 : (b <= 80 * 1024 ? 2 : 1); };
        if (wg(mb) >= wg(pl.lds_bytes)) {
            g = gm;
            pl.lds_bytes = mb;
            pl.merge = 1;
        }
    }
    pl.g = g;
    pl.TZ = TZ;
    pl.Tmax = Tmax;
    if (pl.lds_bytes > 160 * 1024) return fail(RFBA_EINVAL, "LDS plan %zu bytes exceeds 160 KiB", pl.lds_bytes);
    if (3 * g.NW < TZ) return fail(RFBA_EINVAL, "internal: y-column buffer smaller than TZ rows");
    return RFBA_OK;
}

// Tiles of every item at the phases plan_geometry chose: they partition the
// item's SR box (a tile may reach past it on the low side; the kernel stores
// only voxels inside the box).
void plan_tiles(const std::vector<DevItem> &items, int TZ, Plan &pl) {
    pl.tiles_by_item.clear();
    pl.item_tiles.clear();
    for (size_t i = 0; i < items.size(); i++) {
        const auto &it = items[i];
        const auto &ph = pl.phase[i];
        const int begin = (int)pl.tiles_by_item.size();
        for (int z = it.bz0 - ph[2]; z < it.bz1; z += TZ)
            for (int y = it.by0 - ph[1]; y < it.by1; y += kTY)
                for (int x = it.bx0 - ph[0]; x < it.bx1; x += kTX) pl.tiles_by_item.push_back({(int)i, x, y, z});
        pl.item_tiles.push_back({begin, (int)pl.tiles_by_item.size() - begin});
    }
    // Queue order: most expensive first (frames per tile), then spatial
    // order so that consecutively dequeued tiles share halo data in L2.
    pl.tiles = pl.tiles_by_item;
    std::stable_sort(pl.tiles.begin(), pl.tiles.end(), [&](const DevTile &a, const DevTile &b) {
        const long ca = (long)items[a.item].T, cb = (long)items[b.item].T;
        return ca > cb;
    });
}

int choose_tz(const std::vector<DevItem> &items, const DevSamples &s, const SRRFConf &c, int dtype, Plan &pl,
              int force_tz) {
    if (force_tz == 4 || force_tz == 8 || force_tz == 16) return plan_geometry(items, s, c, dtype, force_tz, pl);
    // Largest voxel column that leaves room for two workgroups per CU, else
    // the largest that fits at all (small magnifications need wide halos).
    // Measured on MI355X: occupancy (workgroups per CU, set by LDS) matters
    // more than the shared y-pass work a taller column saves, so the first
    // shape that fits three workgroups per CU wins, then two, then one.
    Plan p;
    for (int budget : {53 * 1024, 80 * 1024, 160 * 1024})
        for (int tz : {8, 16, 4})
            if (plan_geometry(items, s, c, dtype, tz, p) == RFBA_OK && p.lds_bytes <= (size_t)budget) {
                pl = p;
                return RFBA_OK;
            }
    return plan_geometry(items, s, c, dtype, 4, pl);
}

typedef void (*KernelFn)(LaunchParams);

template <bool RP2> KernelFn pick_kernel_r(int dtype, int TZ) {
    if (dtype == RFBA_U16)
        return TZ == 16 ? srrf_fused<uint16_t, 16, RP2>
                        : (TZ == 8 ? srrf_fused<uint16_t, 8, RP2> : srrf_fused<uint16_t, 4, RP2>);
    return TZ == 16 ? srrf_fused<float, 16, RP2> : (TZ == 8 ? srrf_fused<float, 8, RP2> : srrf_fused<float, 4, RP2>);
}

KernelFn pick_kernel(int dtype, int TZ, const DevSamples &s) {
    return s.r_pow2 ? pick_kernel_r<true>(dtype, TZ) : pick_kernel_r<false>(dtype, TZ);
}

int occupancy_grid(KernelFn fn, size_t lds, int ntiles, int &grid, int ncu = 0) {
    if (lds > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, kThreads, lds));
    if (per_cu <= 0) return fail(RFBA_EINVAL, "kernel does not fit on a CU (LDS %zu)", lds);
    grid = std::max(1, std::min(ntiles, per_cu * (ncu > 0 ? ncu : g_ctx.ncu)));
    return RFBA_OK;
}

LaunchParams make_params(const SRRFConf &c, const Plan &pl) {
    LaunchParams P;
    memset(&P, 0, sizeof P);
    P.method = c.Method;
    P.itc = c.ITC;
    P.order = c.TRACOrder;
    P.remove_con = c.RemoveCon;
    P.renorm = c.Renormalize;
    P.smooth = c.GradientSmooth == 0 ? 0 : 1;
    P.iw = c.IntensityW;
    P.mR = c.MainRange;
    P.xR = c.XRange;
    P.yR = c.YRange;
    P.zR = c.ZRange;
    P.g = pl.g;
    P.merge = pl.merge;
    P.Tmax = pl.Tmax;
    {
        // terms of the widest gradient sum; values below 2^16 keep every
        // partial sum of at most 256 terms below 2^24 (exact in FP32)
        const long long m2 = 2LL * c.MainRange, x1 = 2LL * c.XRange + 1, y1 = 2LL * c.YRange + 1,
                        z1 = 2LL * c.ZRange + 1;
        const long long terms = std::max(m2 * y1 * z1, std::max(m2 * x1 * z1, m2 * x1 * y1));
        const LaunchGeom &g = pl.g;
        const long long px = g.pxv ? 3LL * g.NV * g.VS : 3LL * g.NV * g.NW * g.TXP;
        const long long need = (long long)g.RU * g.RV * g.NW + (long long)g.RU * g.NV * g.RW;
        P.sep = (c.GradientSmooth != 0 && terms <= 256 && need <= px) ? 1 : 0;
    }
    return P;
}

bool needs_series(const SRRFConf &c) { return c.Method == 2 || c.Method == 3; }

// Device buffers one run of the fused kernel uses besides its inputs and
// outputs: the sample tables, the queue counter and error flag, the item
// and tile tables, and the per-workgroup frame series of TRAC / TRPPM.
// Either the context's cached allocations (grown on demand) or fixed
// regions of the library's static scratch array (never grown: a series
// area too small for the occupancy grid reduces the grid instead).
struct RunBufs {
    bool fixed = false;
    int ncu = 0;  // CUs of the current device (0: the context's)
    DevSamples *smp = nullptr;
    int *ctl = nullptr;
    DevItem *items = nullptr;
    size_t items_cap = 0;
    DevTile *tiles = nullptr;
    size_t tiles_cap = 0;
    float *series = nullptr;
    size_t series_bytes = 0;
};

// Run items through the fused kernel (single queue).  Synchronous.
int run_items_with(RunBufs &B, const std::vector<DevItem> &items, int dtype, const SRRFConf &c,
                   const DevSamples &s, hipStream_t stream, float *ms) {
    if (items.empty()) return RFBA_OK;
    Plan pl;
    const char *tz_env = getenv("RFBA_SRRF_TZ");  // tuning override: 8 or 16
    int rc = choose_tz(items, s, c, dtype, pl, tz_env ? atoi(tz_env) : 0);
    if (rc) return rc;
    plan_tiles(items, pl.TZ, pl);
    if (pl.tiles.empty()) return RFBA_OK;
    KernelFn fn = pick_kernel(dtype, pl.TZ, s);
    int grid = 0;
    if ((rc = occupancy_grid(fn, pl.lds_bytes, (int)pl.tiles.size(), grid, B.ncu))) return rc;
    const size_t items_bytes = items.size() * sizeof(DevItem), tiles_bytes = pl.tiles.size() * sizeof(DevTile);
    if (B.fixed) {
        if (items_bytes > B.items_cap) return fail(RFBA_EINVAL, "%zu items exceed the scratch table", items.size());
        if (tiles_bytes > B.tiles_cap)
            return fail(RFBA_EINVAL, "%zu SR tiles exceed the scratch tile table of %zu", pl.tiles.size(),
                        B.tiles_cap / sizeof(DevTile));
    } else {
        if ((rc = ensure(g_ctx.d_items, g_ctx.items_cap, items_bytes))) return rc;
        if ((rc = ensure(g_ctx.d_tiles, g_ctx.tiles_cap, tiles_bytes))) return rc;
        B.items = g_ctx.d_items;
        B.tiles = g_ctx.d_tiles;
    }
    LaunchParams P = make_params(c, pl);
    if (needs_series(c)) {
        const size_t stride = (size_t)pl.Tmax * pl.TZ * kThreads;
        if (B.fixed) {
            const size_t fit = B.series_bytes / (stride * sizeof(float));
            if (fit < 1)
                return fail(RFBA_EINVAL, "a frame series of %d frames does not fit the scratch series area",
                            pl.Tmax);
            grid = (int)std::min<size_t>((size_t)grid, fit);
        } else {
            if ((rc = ensure(g_ctx.d_scratch, g_ctx.scratch_bytes, stride * grid * sizeof(float)))) return rc;
            B.series = g_ctx.d_scratch;
        }
        P.scratch = B.series;
        P.scratch_stride = (long long)stride;
    }
    HIP_TRY(hipMemcpyAsync(B.smp, &s, sizeof(DevSamples), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(B.items, items.data(), items_bytes, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(B.tiles, pl.tiles.data(), tiles_bytes, hipMemcpyHostToDevice, stream));
    P.smp = B.smp;
    P.items = B.items;
    P.tiles = B.tiles;
    P.ntiles = (int)pl.tiles.size();
    P.counter = B.ctl;
    P.err = B.ctl + 1;
    if (ms) HIP_TRY(hipEventRecord(g_ctx.ev0, stream));
    hipLaunchKernelGGL(srrf_reset, dim3(1), dim3(64), 0, stream, P.counter, P.err);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(fn, dim3(grid), dim3(kThreads), pl.lds_bytes, stream, P);
    HIP_TRY(hipGetLastError());
    if (ms) HIP_TRY(hipEventRecord(g_ctx.ev1, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    int ctl[2] = {0, 0};
    HIP_TRY(hipMemcpy(ctl, B.ctl, sizeof ctl, hipMemcpyDeviceToHost));
    if (ctl[1]) return fail(RFBA_EDEVICE, "tile geometry exceeded the planned LDS extents");
    if (ms) HIP_TRY(hipEventElapsedTime(ms, g_ctx.ev0, g_ctx.ev1));
    return RFBA_OK;
}

int run_items(const std::vector<DevItem> &items, int dtype, const SRRFConf &c, const DevSamples &s,
              hipStream_t stream, float *ms) {
    RunBufs B;
    B.smp = g_ctx.d_smp;
    B.ctl = g_ctx.d_ctl;
    return run_items_with(B, items, dtype, c, s, stream, ms);
}

DevItem stack_item(const void *d_in, const int dims[4], int M, float *d_out) {
    DevItem it;
    memset(&it, 0, sizeof it);
    it.src = d_in;
    it.sy = dims[0];
    it.sz = (long long)dims[0] * dims[1];
    it.st = (long long)dims[0] * dims[1] * dims[2];
    it.W = dims[0]; it.H = dims[1]; it.D = dims[2]; it.T = dims[3];
    it.M = M;
    it.out = d_out;
    it.osy = (long long)dims[0] * M;
    it.osz = (long long)dims[0] * M * dims[1] * M;
    it.bx0 = it.by0 = it.bz0 = 0;
    it.bx1 = dims[0] * M; it.by1 = dims[1] * M; it.bz1 = dims[2] * M;
    return it;
}

int check_dims(const int *dims) {
    if (!dims) return fail(RFBA_EINVAL, "dims is NULL");
    for (int i = 0; i < 4; i++)
        if (dims[i] <= 0) return fail(RFBA_EINVAL, "dimension %d is %d", i, dims[i]);
    return RFBA_OK;
}

size_t dtype_size(int dtype) { return dtype == RFBA_U16 ? 2 : (dtype == RFBA_F32 ? 4 : 0); }

int stack_device_locked(const void *d_in, int dtype, const int dims[4], const SRRFConf *conf,
                        const rfba_srrf_samples *smp, float *d_out, hipStream_t stream, float *ms) {
    int rc;
    if ((rc = check_conf(conf)) || (rc = check_dims(dims))) return rc;
    if (!dtype_size(dtype)) return fail(RFBA_EINVAL, "unknown dtype %d", dtype);
    if (!d_in || !d_out) return fail(RFBA_EINVAL, "NULL buffer");
    if ((rc = ctx_init())) return rc;
    DevSamples s;
    if ((rc = samples_from(conf, smp, s))) return rc;
    std::vector<DevItem> items{stack_item(d_in, dims, conf->Magnification, d_out)};
    return run_items(items, dtype, *conf, s, stream, ms);
}

int stack_host_locked(const void *in, int dtype, const int dims[4], const SRRFConf *conf,
                      const rfba_srrf_samples *smp, float *out) {
    int rc;
    if ((rc = check_conf(conf)) || (rc = check_dims(dims))) return rc;
    if (!dtype_size(dtype)) return fail(RFBA_EINVAL, "unknown dtype %d", dtype);
    if (!in || !out) return fail(RFBA_EINVAL, "NULL buffer");
    if ((rc = ctx_init())) return rc;
    const size_t M = conf->Magnification;
    const size_t nin = (size_t)dims[0] * dims[1] * dims[2] * dims[3] * dtype_size(dtype);
    const size_t nout = (size_t)dims[0] * dims[1] * dims[2] * M * M * M * sizeof(float);
    if ((rc = ensure(g_ctx.d_in, g_ctx.in_bytes, nin))) return rc;
    if ((rc = ensure(g_ctx.d_out, g_ctx.out_bytes, nout))) return rc;
    HIP_TRY(hipMemcpy(g_ctx.d_in, in, nin, hipMemcpyHostToDevice));
    if ((rc = stack_device_locked(g_ctx.d_in, dtype, dims, conf, smp, g_ctx.d_out, nullptr, nullptr))) return rc;
    HIP_TRY(hipMemcpy(out, g_ctx.d_out, nout, hipMemcpyDeviceToHost));
    return RFBA_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" int rfba_srrf_stack(const void *in, int dtype, const int dims[4], const SRRFConf *conf,
                               const rfba_srrf_samples *samples, float *out) {
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    return stack_host_locked(in, dtype, dims, conf, samples, out);
}

extern "C" int rfba_srrf_stack_device(const void *d_in, int dtype, const int dims[4], const SRRFConf *conf,
                                      const rfba_srrf_samples *samples, float *d_out, void *stream,
                                      float *kernel_ms) {
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    return stack_device_locked(d_in, dtype, dims, conf, samples, d_out, (hipStream_t)stream, kernel_ms);
}

// ---------------------------------------------------------------------------
// Allocation-free form of rfba_srrf_stack_device.  Everything the run needs
// besides the caller's input and output lives in one file-scope device array
// (each GPU holds its own copy until the process ends):
//   [0, 256)           queue counter and error flag
//   [256, 64 KiB+256)  sample tables (DevSamples)
//   next 256 bytes     the stack's item
//   next 16 MiB        SR tile table (up to 1,048,576 tiles)
//   the rest           TRAC / TRPPM frame series, one slot per workgroup
// Host code never names the array directly: its address on each device comes
// from hipGetSymbolAddress at the first call with that device current.

constexpr size_t kNaBytes = (size_t)256 << 20;
constexpr size_t kNaCtl = 256, kNaSmp = 64 * 1024, kNaItem = 256, kNaTiles = (size_t)16 << 20;
constexpr size_t kNaSeries = kNaBytes - kNaCtl - kNaSmp - kNaItem - kNaTiles;
static_assert(sizeof(DevSamples) <= kNaSmp, "sample tables exceed their scratch region");
static_assert(sizeof(DevItem) <= kNaItem, "item exceeds its scratch region");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
constexpr int kNaMaxDev = 64;
struct NaDevice {
    bool ready = false;
    int ncu = 0;
    unsigned char *ptr = nullptr;  // device address of this device's scratch
};
NaDevice g_na[kNaMaxDev];

int noalloc_device(NaDevice *&out) {
    int n = 0, dev = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(RFBA_ENODEV, "no HIP device");
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDev) return fail(RFBA_ENODEV, "device %d beyond the %d supported", dev, kNaMaxDev);
    NaDevice &d = g_na[dev];
    if (!d.ready) {
        hipDeviceProp_t prop;
        HIP_TRY(hipGetDeviceProperties(&prop, dev));
        if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            return fail(RFBA_ENODEV, "device %s is not gfx950", prop.gcnArchName);
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(RFBA_EDEVICE, "scratch address: %s", hipGetErrorString(e));
        d.ncu = prop.multiProcessorCount;
        d.ptr = (unsigned char *)p;
        d.ready = true;
    }
    out = &d;
    return RFBA_OK;
}
}  // namespace

extern "C" int rfba_srrf_stack_device_noalloc(const void *d_in, int dtype, const int dims[4], const SRRFConf *conf,
                                              const rfba_srrf_samples *samples, float *d_out, void *stream) {
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    int rc;
    if ((rc = check_conf(conf)) || (rc = check_dims(dims))) return rc;
    if (!dtype_size(dtype)) return fail(RFBA_EINVAL, "unknown dtype %d", dtype);
    if (!d_in || !d_out) return fail(RFBA_EINVAL, "NULL buffer");
    NaDevice *nd = nullptr;
    if ((rc = noalloc_device(nd))) return rc;
    DevSamples s;
    if ((rc = samples_from(conf, samples, s))) return rc;
    RunBufs B;
    B.fixed = true;
    B.ncu = nd->ncu;
    unsigned char *base = nd->ptr;
    B.ctl = (int *)base;
    B.smp = (DevSamples *)(base + kNaCtl);
    B.items = (DevItem *)(base + kNaCtl + kNaSmp);
    B.items_cap = kNaItem;
    B.tiles = (DevTile *)(base + kNaCtl + kNaSmp + kNaItem);
    B.tiles_cap = kNaTiles;
    B.series = (float *)(base + kNaCtl + kNaSmp + kNaItem + kNaTiles);
    B.series_bytes = kNaSeries;
    std::vector<DevItem> items{stack_item(d_in, dims, conf->Magnification, d_out)};
    return run_items_with(B, items, dtype, *conf, s, (hipStream_t)stream, nullptr);
}

extern "C" int SRRFGradientCalcGPU(float *inImage, int *imageDim, float *outImage, SRRFConf *conf) {
    // kernel.cu:560-657: one stack, reference ring sampling, host buffers.
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    return stack_host_locked(inImage, RFBA_F32, imageDim, conf, nullptr, outImage);
}

extern "C" int SRRFGetMag(void) {
    // GradientCalc.cpp:176-178
    Ini ini;
    ini.load("SRRF.ini");
    return ini.get_int("SRRF", "Magnification", 5);
}

// Tiles of SRRFGradientCalc (GradientCalc.cpp:47-116) as batched items.  Each
// tile is its own stack (own clamp-to-edge borders and NaN bands).  The
// reference copies tile voxels [2M, res*M) per axis into the output in loop
// order k, j, i, later tiles overwriting earlier ones, so the surviving part
// of tile p along an axis is [2M, (B-2)M) (last tile: [2M, res*M)); only that
// box is computed and written, straight into the output.
static int tiled_items(const int *imageDim, const SRRFConf &c, const void *d_in, float *d_out,
                       std::vector<DevItem> &items) {
    const int border = 2, M = c.Magnification;
    const int maxDim[3] = {c.blockSizeX, c.blockSizeY, c.blockSizeZ};
    int splitN[3];
    for (int i = 0; i < 3; i++) {
        if (maxDim[i] - 2 * border <= 0) return fail(RFBA_EINVAL, "blockSize %d <= 4", maxDim[i]);
        splitN[i] = (imageDim[i] - 2 * border) / (maxDim[i] - 2 * border);
        if ((imageDim[i] - 2 * border) % (maxDim[i] - 2 * border) != 0) splitN[i] = splitN[i] + 1;
    }
    const long long OX = (long long)imageDim[0] * M, OY = (long long)imageDim[1] * M;
    for (int k = 0; k < splitN[2]; k++)
        for (int j = 0; j < splitN[1]; j++)
            for (int i = 0; i < splitN[0]; i++) {
                const int posN[3] = {i, j, k};
                int res[3], start[3], hi[3];
                for (int a = 0; a < 3; a++) {
                    const bool last = posN[a] == splitN[a] - 1;
                    res[a] = last ? imageDim[a] - (splitN[a] - 1) * (maxDim[a] - 2 * border) : maxDim[a];
                    start[a] = posN[a] == 0 ? 0 : posN[a] * (maxDim[a] - 2 * border);
                    hi[a] = last ? res[a] * M : (maxDim[a] - 2) * M;
                    hi[a] = std::min(hi[a], res[a] * M);
                }
                DevItem it;
                memset(&it, 0, sizeof it);
                it.src = d_in;
                it.sy = imageDim[0];
                it.sz = (long long)imageDim[0] * imageDim[1];
                it.st = (long long)imageDim[0] * imageDim[1] * imageDim[2];
                it.ox = start[0]; it.oy = start[1]; it.oz = start[2];
                it.W = res[0]; it.H = res[1]; it.D = res[2]; it.T = imageDim[3];
                it.M = M;
                it.osy = OX;
                it.osz = OX * OY;
                it.out = d_out + (long long)start[2] * M * it.osz + (long long)start[1] * M * it.osy +
                         (long long)start[0] * M;
                it.bx0 = it.by0 = it.bz0 = 2 * M;
                it.bx1 = hi[0]; it.by1 = hi[1]; it.bz1 = hi[2];
                if (it.bx1 > it.bx0 && it.by1 > it.by0 && it.bz1 > it.bz0) items.push_back(it);
            }
    return RFBA_OK;
}

extern "C" int SRRFGradientCalc(float *inImage, int *imageDim, float *outImage) {
    // GradientCalc.cpp:169-173 -> doSRRFGradientCalc :120-163
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    SRRFConf c;
    int rc;
    if ((rc = rfba_srrf_load_ini("./SRRF.ini", &c))) return rc;
    if ((rc = check_conf(&c)) || (rc = check_dims(imageDim))) return rc;
    if (!inImage || !outImage) return fail(RFBA_EINVAL, "NULL buffer");
    if ((rc = ctx_init())) return rc;
    const size_t M = c.Magnification;
    const size_t nin = (size_t)imageDim[0] * imageDim[1] * imageDim[2] * imageDim[3] * sizeof(float);
    const size_t nout = (size_t)imageDim[0] * imageDim[1] * imageDim[2] * M * M * M * sizeof(float);
    if ((rc = ensure(g_ctx.d_in, g_ctx.in_bytes, nin))) return rc;
    if ((rc = ensure(g_ctx.d_out, g_ctx.out_bytes, nout))) return rc;
    HIP_TRY(hipMemcpy(g_ctx.d_in, inImage, nin, hipMemcpyHostToDevice));
    // Voxels the reference never writes keep the caller's values.
    HIP_TRY(hipMemcpy(g_ctx.d_out, outImage, nout, hipMemcpyHostToDevice));
    std::vector<DevItem> items;
    if ((rc = tiled_items(imageDim, c, g_ctx.d_in, g_ctx.d_out, items))) return rc;
    DevSamples s;
    if ((rc = samples_from(&c, nullptr, s))) return rc;
    if ((rc = run_items(items, RFBA_F32, c, s, nullptr, nullptr))) return rc;
    HIP_TRY(hipMemcpy(outImage, g_ctx.d_out, nout, hipMemcpyDeviceToHost));
    return RFBA_OK;
}

// ---------------------------------------------------------------------------
// batched form

// Blocks are grouped by magnification: each group has its own tile depth,
// LDS layout, kernel and grid (a mixed batch planned as one would size every
// tile's LDS for the smallest magnification's halo and lose occupancy).  The
// groups' persistent launches run back to back on one stream (one graph).
struct rfba_srrf_batch {
    SRRFConf conf;
    int dtype = 0;
    std::vector<DevItem> items;  // all blocks, caller order
    DevSamples smp;
    DevSamples *d_smp = nullptr;
    DevItem *d_items = nullptr;
    int *d_ctl = nullptr;        // [2 per group]: queue counter, error flag
    float *d_scratch = nullptr;  // TRAC/TRPPM series, shared by the groups' launches
    size_t scratch_bytes = 0;
    struct Group {
        int M = 0;
        std::vector<int> ids;         // global item indices
        std::vector<DevItem> gitems;  // the group's items (planning)
        Plan pl;
        KernelFn fn = nullptr;
        int grid = 0;
        long long scratch_stride = 0;
        DevTile *d_tiles = nullptr;          // queue order
        DevTile *d_tiles_by_item = nullptr;  // per block, for the per-block form
        double work = 0.0;                   // frames x SR voxels
        float tune_ms[3] = {0, 0, 0};        // sample time at TZ 4, 8, 16 (0: not timed)
    };
    std::vector<Group> groups;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

static void batch_drop_graph(rfba_srrf_batch *b) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

static void batch_free(rfba_srrf_batch *b) {
    if (!b) return;
    batch_drop_graph(b);
    for (auto &g : b->groups) {
        if (g.d_tiles) hipFree(g.d_tiles);
        if (g.d_tiles_by_item) hipFree(g.d_tiles_by_item);
    }
    if (b->d_smp) hipFree(b->d_smp);
    if (b->d_items) hipFree(b->d_items);
    if (b->d_ctl) hipFree(b->d_ctl);
    if (b->d_scratch) hipFree(b->d_scratch);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    delete b;
}

static int realloc_dev(void **p, size_t bytes) {
    if (*p) hipFree(*p);
    *p = nullptr;
    HIP_TRY(hipMalloc(p, std::max<size_t>(bytes, 16)));
    return RFBA_OK;
}

// (Re)build one group's device side for the tile depth in geom: tile queue
// (tiles name global item indices), kernel, grid, series scratch.
static int group_build(rfba_srrf_batch *b, rfba_srrf_batch::Group &g, const Plan &geom) {
    int rc;
    batch_drop_graph(b);
    g.pl = geom;
    plan_tiles(g.gitems, g.pl.TZ, g.pl);
    for (auto &t : g.pl.tiles) t.item = g.ids[t.item];
    for (auto &t : g.pl.tiles_by_item) t.item = g.ids[t.item];
    g.fn = pick_kernel(b->dtype, g.pl.TZ, b->smp);
    if ((rc = occupancy_grid(g.fn, g.pl.lds_bytes, (int)g.pl.tiles.size(), g.grid))) return rc;
    const size_t nt = g.pl.tiles.size();
    if ((rc = realloc_dev((void **)&g.d_tiles, nt * sizeof(DevTile))) ||
        (rc = realloc_dev((void **)&g.d_tiles_by_item, nt * sizeof(DevTile))))
        return rc;
    HIP_TRY(hipMemcpy(g.d_tiles, g.pl.tiles.data(), nt * sizeof(DevTile), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(g.d_tiles_by_item, g.pl.tiles_by_item.data(), nt * sizeof(DevTile), hipMemcpyHostToDevice));
    g.scratch_stride = 0;
    if (needs_series(b->conf)) {
        g.scratch_stride = (long long)g.pl.Tmax * g.pl.TZ * kThreads;
        const size_t need = (size_t)g.scratch_stride * g.grid * sizeof(float);
        if (need > b->scratch_bytes) {
            if ((rc = realloc_dev((void **)&b->d_scratch, need))) return rc;
            b->scratch_bytes = need;
        }
    }
    return RFBA_OK;
}

// Device tables shared by the groups: samples, items, per-group control words.
static int batch_tables(rfba_srrf_batch *b) {
    int rc;
    if ((rc = realloc_dev((void **)&b->d_smp, sizeof(DevSamples))) ||
        (rc = realloc_dev((void **)&b->d_items, b->items.size() * sizeof(DevItem))) ||
        (rc = realloc_dev((void **)&b->d_ctl, 2 * b->groups.size() * sizeof(int))))
        return rc;
    HIP_TRY(hipMemcpy(b->d_smp, &b->smp, sizeof b->smp, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->d_items, b->items.data(), b->items.size() * sizeof(DevItem), hipMemcpyHostToDevice));
    return RFBA_OK;
}

// Group the items by magnification, largest work first.
static void batch_group(rfba_srrf_batch *b) {
    b->groups.clear();
    std::vector<int> mags;
    for (const auto &it : b->items)
        if (std::find(mags.begin(), mags.end(), it.M) == mags.end()) mags.push_back(it.M);
    for (int M : mags) {
        rfba_srrf_batch::Group g;
        g.M = M;
        for (size_t i = 0; i < b->items.size(); i++)
            if (b->items[i].M == M) {
                g.ids.push_back((int)i);
                g.gitems.push_back(b->items[i]);
                const DevItem &it = b->items[i];
                g.work += (double)it.T * (double)(it.bx1 - it.bx0) * (it.by1 - it.by0) * (it.bz1 - it.bz0);
            }
        b->groups.push_back(std::move(g));
    }
    std::stable_sort(b->groups.begin(), b->groups.end(),
                     [](const rfba_srrf_batch::Group &x, const rfba_srrf_batch::Group &y) { return x.work > y.work; });
}

extern "C" int rfba_srrf_batch_create(const rfba_srrf_block *blocks, int n, int dtype, const SRRFConf *conf,
                                      const rfba_srrf_samples *samples, rfba_srrf_batch **out) {
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    int rc;
    if (!out) return fail(RFBA_EINVAL, "plan pointer is NULL");
    *out = nullptr;
    if (!blocks || n <= 0) return fail(RFBA_EINVAL, "no blocks");
    if ((rc = check_conf(conf))) return rc;
    if (!dtype_size(dtype)) return fail(RFBA_EINVAL, "unknown dtype %d", dtype);
    if ((rc = ctx_init())) return rc;
    DevSamples s;
    if ((rc = samples_from(conf, samples, s))) return rc;
    rfba_srrf_batch *b = new rfba_srrf_batch();
    b->conf = *conf;
    b->dtype = dtype;
    b->smp = s;
    for (int i = 0; i < n; i++) {
        const rfba_srrf_block &bl = blocks[i];
        for (int a = 0; a < 3; a++)
            if (bl.dims[a] <= 0 || bl.origin[a] < 0 || bl.origin[a] + bl.dims[a] > bl.src_dims[a]) {
                batch_free(b);
                return fail(RFBA_EINVAL, "block %d axis %d outside its source", i, a);
            }
        if (bl.frames <= 0 || bl.frames > bl.src_dims[3] || bl.mag <= 0 || bl.mag > 64 || !bl.d_src || !bl.d_out) {
            batch_free(b);
            return fail(RFBA_EINVAL, "block %d: invalid frames/mag/buffers", i);
        }
        DevItem it;
        memset(&it, 0, sizeof it);
        it.src = bl.d_src;
        it.sy = bl.src_dims[0];
        it.sz = (long long)bl.src_dims[0] * bl.src_dims[1];
        it.st = (long long)bl.src_dims[0] * bl.src_dims[1] * bl.src_dims[2];
        it.ox = bl.origin[0]; it.oy = bl.origin[1]; it.oz = bl.origin[2];
        it.W = bl.dims[0]; it.H = bl.dims[1]; it.D = bl.dims[2]; it.T = bl.frames;
        it.M = bl.mag;
        it.out = bl.d_out;
        it.osy = (long long)bl.dims[0] * bl.mag;
        it.osz = it.osy * bl.dims[1] * bl.mag;
        it.bx1 = bl.dims[0] * bl.mag; it.by1 = bl.dims[1] * bl.mag; it.bz1 = bl.dims[2] * bl.mag;
        b->items.push_back(it);
    }
    batch_group(b);
    for (auto &g : b->groups) {
        Plan p;
        if ((rc = choose_tz(g.gitems, s, *conf, b->dtype, p, 0)) || (rc = group_build(b, g, p))) {
            batch_free(b);
            return rc;
        }
    }
    if ((rc = batch_tables(b))) {
        batch_free(b);
        return rc;
    }
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e != hipSuccess) {
        batch_free(b);
        return fail(RFBA_EDEVICE, "batch setup: %s", hipGetErrorString(e));
    }
    *out = b;
    return RFBA_OK;
}

static LaunchParams batch_params(rfba_srrf_batch *b, const rfba_srrf_batch::Group &g, int gi, const DevTile *tiles,
                                 int ntiles) {
    LaunchParams P = make_params(b->conf, g.pl);
    P.smp = b->d_smp;
    P.items = b->d_items;
    P.tiles = tiles;
    P.ntiles = ntiles;
    P.counter = b->d_ctl + 2 * gi;
    P.err = b->d_ctl + 2 * gi + 1;
    P.scratch = b->d_scratch;
    P.scratch_stride = g.scratch_stride;
    return P;
}

static int batch_enqueue(rfba_srrf_batch *b, hipStream_t st) {
    for (size_t gi = 0; gi < b->groups.size(); gi++) {
        const auto &g = b->groups[gi];
        if (g.pl.tiles.empty()) continue;
        LaunchParams P = batch_params(b, g, (int)gi, g.d_tiles, (int)g.pl.tiles.size());
        hipLaunchKernelGGL(srrf_reset, dim3(1), dim3(64), 0, st, P.counter, P.err);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(g.fn, dim3(g.grid), dim3(kThreads), g.pl.lds_bytes, st, P);
        HIP_TRY(hipGetLastError());
    }
    return RFBA_OK;
}

static int batch_check(rfba_srrf_batch *b) {
    std::vector<int> ctl(2 * b->groups.size());
    HIP_TRY(hipMemcpy(ctl.data(), b->d_ctl, ctl.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t gi = 0; gi < b->groups.size(); gi++)
        if (ctl[2 * gi + 1]) return fail(RFBA_EDEVICE, "tile geometry exceeded the planned LDS extents");
    return RFBA_OK;
}

extern "C" int rfba_srrf_batch_run(rfba_srrf_batch *b, void *stream, int use_graph, float *kernel_ms) {
    if (!b) return fail(RFBA_EINVAL, "plan is NULL");
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if (use_graph) {
        if (!b->exec) {
            hipStream_t cs;
            HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            rc = batch_enqueue(b, cs);
            hipGraph_t gr = nullptr;
            hipError_t e = hipStreamEndCapture(cs, &gr);
            hipStreamDestroy(cs);
            if (rc) return rc;
            if (e != hipSuccess) return fail(RFBA_EDEVICE, "graph capture: %s", hipGetErrorString(e));
            b->graph = gr;
            HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        HIP_TRY(hipEventRecord(b->ev0, st));
        HIP_TRY(hipGraphLaunch(b->exec, st));
        HIP_TRY(hipEventRecord(b->ev1, st));
    } else {
        HIP_TRY(hipEventRecord(b->ev0, st));
        if ((rc = batch_enqueue(b, st))) return rc;
        HIP_TRY(hipEventRecord(b->ev1, st));
    }
    if (kernel_ms) {
        HIP_TRY(hipEventSynchronize(b->ev1));
        HIP_TRY(hipEventElapsedTime(kernel_ms, b->ev0, b->ev1));
        return batch_check(b);
    }
    return RFBA_OK;
}

extern "C" int rfba_srrf_batch_run_per_block(rfba_srrf_batch *b, void *stream, float *kernel_ms) {
    if (!b) return fail(RFBA_EINVAL, "plan is NULL");
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipEventRecord(b->ev0, st));
    for (size_t gi = 0; gi < b->groups.size(); gi++) {
        const auto &g = b->groups[gi];
        for (size_t i = 0; i < g.pl.item_tiles.size(); i++) {
            const auto &r = g.pl.item_tiles[i];
            if (r.second == 0) continue;
            LaunchParams P = batch_params(b, g, (int)gi, g.d_tiles_by_item + r.first, r.second);
            const int grid = std::min(g.grid, r.second);
            hipLaunchKernelGGL(srrf_reset, dim3(1), dim3(64), 0, st, P.counter, P.err);
            HIP_TRY(hipGetLastError());
            hipLaunchKernelGGL(g.fn, dim3(grid), dim3(kThreads), g.pl.lds_bytes, st, P);
            HIP_TRY(hipGetLastError());
        }
    }
    HIP_TRY(hipEventRecord(b->ev1, st));
    if (kernel_ms) {
        HIP_TRY(hipEventSynchronize(b->ev1));
        HIP_TRY(hipEventElapsedTime(kernel_ms, b->ev0, b->ev1));
        return batch_check(b);
    }
    return RFBA_OK;
}

extern "C" int rfba_srrf_batch_status(rfba_srrf_batch *b) {
    if (!b) return fail(RFBA_EINVAL, "plan is NULL");
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    HIP_TRY(hipEventSynchronize(b->ev1));
    return batch_check(b);
}

// Autotuner: for each magnification group, each feasible tile depth
// (geometry planned over all of the group's blocks, so that the LDS layout
// is the one the full run uses) is timed on a one-group sub-plan of every
// sample_stride-th block of the group, one warm run then one timed run; the
// fastest depth is kept and the group rebuilt for it.  The sampled blocks'
// outputs are written (with their correct values).
extern "C" int rfba_srrf_batch_autotune(rfba_srrf_batch *b, void *stream, int sample_stride, int *tz,
                                        float *ms_by_tz) {
    if (!b) return fail(RFBA_EINVAL, "plan is NULL");
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    hipStream_t st = (hipStream_t)stream;
    const int cand[3] = {4, 8, 16};
    int rc = RFBA_OK;
    for (auto &g : b->groups) {
        const int n = (int)g.ids.size();
        const int stride = sample_stride > 0 ? sample_stride : std::max(1, n / 32);
        rfba_srrf_batch *sub = new rfba_srrf_batch();
        sub->conf = b->conf;
        sub->dtype = b->dtype;
        sub->smp = b->smp;
        rfba_srrf_batch::Group sg;
        sg.M = g.M;
        for (int i = 0; i < n; i += stride) {
            sg.ids.push_back((int)sub->items.size());
            sg.gitems.push_back(g.gitems[i]);
            sub->items.push_back(g.gitems[i]);
        }
        sub->groups.push_back(std::move(sg));
        Plan best;
        float best_ms = 0.0f;
        if ((rc = batch_tables(sub))) {
            batch_free(sub);
            return rc;
        }
        for (int c = 0; c < 3 && rc == RFBA_OK; c++) {
            g.tune_ms[c] = 0.0f;
            Plan p;
            if (plan_geometry(g.gitems, b->smp, b->conf, b->dtype, cand[c], p) != RFBA_OK) continue;
            if ((rc = group_build(sub, sub->groups[0], p))) break;
            float ms = 0.0f;
            for (int rep = 0; rep < 2 && rc == RFBA_OK; rep++) {
                if ((rc = hip_status(hipEventRecord(b->ev0, st)))) break;
                if ((rc = batch_enqueue(sub, st))) break;
                if ((rc = hip_status(hipEventRecord(b->ev1, st)))) break;
                if ((rc = hip_status(hipEventSynchronize(b->ev1)))) break;
                if ((rc = hip_status(hipEventElapsedTime(&ms, b->ev0, b->ev1)))) break;
                rc = batch_check(sub);
            }
            if (rc) break;
            g.tune_ms[c] = ms;
            if (best_ms == 0.0f || ms < best_ms) {
                best_ms = ms;
                best = p;
            }
        }
        batch_free(sub);
        if (rc) return rc;
        if (best_ms == 0.0f) return fail(RFBA_EINVAL, "no feasible tile depth for magnification %d", g.M);
        if ((rc = group_build(b, g, best))) return rc;
    }
    if (!b->groups.empty()) {
        const auto &g0 = b->groups[0];  // the group with the most work
        if (tz) *tz = g0.pl.TZ;
        if (ms_by_tz)
            for (int c = 0; c < 3; c++) ms_by_tz[c] = g0.tune_ms[c];
    }
    return RFBA_OK;
}

extern "C" int rfba_srrf_batch_tile_depth(const rfba_srrf_batch *b) {
    return (b && !b->groups.empty()) ? b->groups[0].pl.TZ : 0;
}

extern "C" int rfba_srrf_batch_groups(const rfba_srrf_batch *b, int *mags, int *tile_depths, int max_groups) {
    if (!b) return fail(RFBA_EINVAL, "plan is NULL");
    const int n = (int)b->groups.size();
    for (int i = 0; i < n && i < max_groups; i++) {
        if (mags) mags[i] = b->groups[i].M;
        if (tile_depths) tile_depths[i] = b->groups[i].pl.TZ;
    }
    return n;
}

extern "C" void rfba_srrf_batch_destroy(rfba_srrf_batch *b) {
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    batch_free(b);
}

// ---------------------------------------------------------------------------
// runtime helpers

extern "C" int rfba_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
extern "C" int rfba_malloc_device(void **p, size_t bytes) {
    if (!p) return fail(RFBA_EINVAL, "NULL");
    HIP_TRY(hipMalloc(p, bytes));
    return RFBA_OK;
}
extern "C" int rfba_free_device(void *p) {
    HIP_TRY(hipFree(p));
    return RFBA_OK;
}
extern "C" int rfba_memcpy_h2d(void *d, const void *s, size_t n) {
    HIP_TRY(hipMemcpy(d, s, n, hipMemcpyHostToDevice));
    return RFBA_OK;
}
extern "C" int rfba_memcpy_d2h(void *d, const void *s, size_t n) {
    HIP_TRY(hipMemcpy(d, s, n, hipMemcpyDeviceToHost));
    return RFBA_OK;
}
extern "C" int rfba_memset_device(void *d, int v, size_t n) {
    HIP_TRY(hipMemset(d, v, n));
    return RFBA_OK;
}
extern "C" int rfba_device_synchronize(void) {
    HIP_TRY(hipDeviceSynchronize());
    return RFBA_OK;
}
extern "C" void rfba_release_cache(void) {
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    auto rel = [](void *&p, size_t &cap) {
        if (p) hipFree(p);
        p = nullptr;
        cap = 0;
    };
    rel((void *&)g_ctx.d_scratch, g_ctx.scratch_bytes);
    rel(g_ctx.d_in, g_ctx.in_bytes);
    rel((void *&)g_ctx.d_out, g_ctx.out_bytes);
    rel((void *&)g_ctx.d_items, g_ctx.items_cap);
    rel((void *&)g_ctx.d_tiles, g_ctx.tiles_cap);
}

// ---------------------------------------------------------------------------
// caller post-processing and the SRRF.raw file (multispot5_place_methods.cc:
// 260-274 write, 284-289 read)

extern "C" int rfba_srrf_postprocess_device(float *d_out, size_t n, void *stream, float *fmax) {
    if (!d_out && n) return fail(RFBA_EINVAL, "NULL buffer");
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    int rc;
    if ((rc = ctx_init())) return rc;
    hipStream_t st = (hipStream_t)stream;
    unsigned *d_max = nullptr;
    HIP_TRY(hipMalloc(&d_max, sizeof(unsigned)));
    hipLaunchKernelGGL(srrf_clamp_init, dim3(1), dim3(64), 0, st, d_max);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && n) {
        const size_t blocks = std::min<size_t>((n + 255) / 256, (size_t)g_ctx.ncu * 16);
        hipLaunchKernelGGL(srrf_clamp, dim3((unsigned)blocks), dim3(256), 0, st, d_out, n, d_max);
        e = hipGetLastError();
    }
    unsigned bits = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&bits, d_max, sizeof bits, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFree(d_max);
    if (e != hipSuccess) return fail(RFBA_EDEVICE, "post-processing: %s", hipGetErrorString(e));
    if (fmax) memcpy(fmax, &bits, sizeof bits);
    return RFBA_OK;
}

extern "C" int rfba_srrf_write_raw(const char *path, const float *data, size_t n) {
    if (!path || (!data && n)) return fail(RFBA_EINVAL, "NULL argument");
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail(RFBA_EINVAL, "cannot open %s for writing", path);
    const size_t w = n ? fwrite(data, sizeof(float), n, fp) : 0;
    const int c = fclose(fp);
    if (w != n || c != 0) return fail(RFBA_EINVAL, "short write to %s", path);
    return RFBA_OK;
}

extern "C" int rfba_srrf_read_raw(const char *path, float *data, size_t n) {
    if (!path || (!data && n)) return fail(RFBA_EINVAL, "NULL argument");
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail(RFBA_EINVAL, "cannot open %s", path);
    const size_t r = n ? fread(data, sizeof(float), n, fp) : 0;
    fclose(fp);
    if (r != n) return fail(RFBA_EINVAL, "%s holds %zu of %zu floats", path, r, n);
    return RFBA_OK;
}

// ---------------------------------------------------------------------------
// SRRF-guided spot placement (place_spots_intensity_sampled,
// multispot5_place_methods.cc:293-341): the region's SR values are gathered
// on the device; the rejection sampling runs on the host with the
// reference's generator (MT19937 seeded as get_defaults :69-71 does, 37
// outputs discarded by RandomInit, draws = output * 2^-32).

extern "C" int rfba_srrf_place_spots(const float *d_sr, const int *dims, int mag, const int *pixels, int npix,
                                     int num_spots, double intensity_mode, double blur_mode, int seed, double *spots,
                                     unsigned long long *draws_used) {
    if (!d_sr || !dims || !pixels || npix <= 0 || mag <= 0 || num_spots < 0 || (num_spots && !spots))
        return fail(RFBA_EINVAL, "invalid arguments");
    const int X = dims[0], Y = dims[1], Z = dims[2];
    for (int i = 0; i < npix; i++)
        if (pixels[3 * i] < 0 || pixels[3 * i] >= X || pixels[3 * i + 1] < 0 || pixels[3 * i + 1] >= Y ||
            pixels[3 * i + 2] < 0 || pixels[3 * i + 2] >= Z)
            return fail(RFBA_EINVAL, "pixel %d outside the volume", i);
    const long long n = (long long)npix * mag * mag * mag;
    std::vector<float> val((size_t)n);
    {
        std::lock_guard<std::mutex> lk(g_ctx.mu);
        int rc;
        if ((rc = ctx_init())) return rc;
        int *d_px = nullptr;
        float *d_val = nullptr;
        hipError_t e = hipMalloc(&d_px, (size_t)npix * 3 * sizeof(int));
        if (e == hipSuccess) e = hipMalloc(&d_val, (size_t)n * sizeof(float));
        if (e == hipSuccess) e = hipMemcpy(d_px, pixels, (size_t)npix * 3 * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            const long long blocks = std::min<long long>((n + 255) / 256, (long long)g_ctx.ncu * 16);
            hipLaunchKernelGGL(srrf_gather_region, dim3((unsigned)blocks), dim3(256), 0, nullptr, d_sr, X, Y, mag, d_px,
                               n, d_val);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpy(val.data(), d_val, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
        hipFree(d_px);
        hipFree(d_val);
        if (e != hipSuccess) return fail(RFBA_EDEVICE, "region gather: %s", hipGetErrorString(e));
    }
    double mx = val[0];
    for (long long i = 1; i < n; i++) mx = std::max(mx, (double)val[(size_t)i]);
    std::mt19937 g((uint32_t)seed);
    g.discard(37);
    unsigned long long used = 0;
    auto rng = [&]() {
        used++;
        return (double)g() * (1.0 / (65536.0 * 65536.0));
    };
    int placed = 0;
    if (!(mx < 0)) {
        const int M3 = mag * mag * mag;
        while (placed < num_spots) {
            const long long element = (long long)std::floor(rng() * (double)n);
            const double yv = rng() * mx;
            if (yv <= (double)val[(size_t)element]) {
                const long long p = element / M3;
                const int j = (int)(element - p * M3);
                const int xs = pixels[3 * p] * mag + j / (mag * mag);
                const int ys = pixels[3 * p + 1] * mag + (j / mag) % mag;
                const int zs = pixels[3 * p + 2] * mag + j % mag;
                float x = (float)(xs + rng() - .5);
                float y = (float)(ys + rng() - .5);
                float z = (float)(zs + rng() - .5);
                x = x / mag;
                y = y / mag;
                z = z / mag;
                double *sp = spots + 5 * (size_t)placed;
                sp[0] = intensity_mode;
                sp[1] = blur_mode;
                sp[2] = x;
                sp[3] = y;
                sp[4] = z;
                placed++;
            }
        }
    }
    if (draws_used) *draws_used = used;
    return placed;
}

extern "C" int rfba_copy_probe(void *d_dst, const void *d_src, size_t bytes, float *ms) {
    if (!d_dst || !d_src || bytes % 16) return fail(RFBA_EINVAL, "copy probe needs 16-byte multiples");
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    int rc;
    if ((rc = ctx_init())) return rc;
    HIP_TRY(hipEventRecord(g_ctx.ev0, nullptr));
    hipLaunchKernelGGL(copy_probe, dim3(g_ctx.ncu * 8), dim3(256), 0, nullptr, (const float4 *)d_src,
                       (float4 *)d_dst, bytes / 16);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(g_ctx.ev1, nullptr));
    HIP_TRY(hipEventSynchronize(g_ctx.ev1));
    if (ms) HIP_TRY(hipEventElapsedTime(ms, g_ctx.ev0, g_ctx.ev1));
    return RFBA_OK;
}
