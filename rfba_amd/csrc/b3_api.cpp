// b3_api.cpp -- host side of librfba_3b.so: the C ABI of include/rfba_3b.h,
// work planning (residual planes, item queue, memory chunks), launches and
// hipGraph capture.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <random>
#include <vector>

#include "b3_internal.h"

namespace rfba3b {
__global__ void b3_reset(int *counter);
__global__ void b3_resid(const DevBlock *__restrict__ blocks, const PlaneTask *__restrict__ tasks, int ntasks,
                         float *__restrict__ R, double *__restrict__ S0, int s0_stride,
                         const DevModel *__restrict__ mdl);
template <int MODE> __global__ void b3_eval(EvalParams E);
template <bool GRAD, bool FUNC, int NCU> __global__ void b3_corr_mfma(CorrParams C);
template <int MODE> __global__ void b3_hmm(HmmParams H);
template <int MODE> __global__ void b3_hmm_wave(HmmParams H);
struct LogA {
    double v[9];
};
__global__ void b3_backward_sample(const double *__restrict__ delta, int n, int per, int T, LogA lA,
                                   const double *__restrict__ u, uint8_t *__restrict__ states);
__global__ void b3_hess_basis(HessParams Hp);
__global__ void b3_corr_reduce(CorrReduceParams R);
__global__ void b3_hess_parts(HessParams Hp);
}  // namespace rfba3b

using namespace rfba3b;

static thread_local char g_err[512] = "";

static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(e_ == hipErrorOutOfMemory ? -2 : -3, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

extern "C" const char *rfba3b_last_error(void) { return g_err; }

namespace {

std::mutex g_mu;
int g_ncu = 0;

int dev_init() {
    if (g_ncu) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(-4, "no HIP device");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return fail(-4, "device %s is not gfx950", prop.gcnArchName);
    g_ncu = prop.multiProcessorCount;
    return 0;
}

int check_model(const rfba3b_model *m) {
    if (!m) return fail(-1, "model is NULL");
    if (!(m->variance > 0) || !std::isfinite(m->variance)) return fail(-1, "variance must be positive");
    if (!(m->zfactor > 0) || !std::isfinite(m->zfactor)) return fail(-1, "zfactor must be positive");
    for (int i = 0; i < 9; i++)
        if (!(m->A[i] >= 0)) return fail(-1, "transition matrix has a negative or NaN entry");
    for (int i = 0; i < 3; i++)
        if (!(m->pi[i] >= 0)) return fail(-1, "initial distribution has a negative or NaN entry");
    if (m->checkpoint < 0) return fail(-1, "checkpoint stride must be >= 0");
    return 0;
}

DevModel dev_model(const rfba3b_model *m) {
    DevModel d;
    memset(&d, 0, sizeof d);
    for (int i = 0; i < 9; i++) d.A[i] = m->A[i];
    for (int i = 0; i < 3; i++) d.pi[i] = m->pi[i];
    d.variance = m->variance;
    d.zfactor = m->zfactor;
    d.z_exact = m->z_exact;
    return d;
}

int ckpt_of(const rfba3b_model *m, int T) {
    return m->checkpoint > 0 ? std::min(m->checkpoint, T) : std::max(1, (int)std::ceil(std::sqrt((double)T)));
}

size_t eval_lds(int Tmax, int XYmax, int Zmax, int ckptmax) {
    const int nseg = (Tmax + 0) / 1 + 1;  // bound: at most Tmax segments (ckpt >= 1)
    // checkpoints (m, e) x 3 per segment; alpha_t(on) and c_t (m, e) per frame of a segment
    return 8 * 8 * (size_t)Tmax + 16 * (size_t)XYmax + 16 * (size_t)Zmax + 16 +
           8 * (6 * (size_t)nseg + 4 * (size_t)ckptmax);
}

typedef void (*EvalFn)(EvalParams);
EvalFn eval_fn(int mode) {
    return mode == MODE_GRAD ? b3_eval<MODE_GRAD> : (mode == MODE_DELTA ? b3_eval<MODE_DELTA> : b3_eval<MODE_LOGL>);
}

int eval_grid(EvalFn fn, size_t lds, int nitems, int &grid) {
    if (lds > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, 256, lds));
    if (per_cu <= 0) return fail(-1, "likelihood kernel does not fit on a CU (LDS %zu bytes)", lds);
    grid = std::max(1, std::min(nitems, per_cu * g_ncu));
    return 0;
}

}  // namespace

namespace rfba3b {
int host_fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}
int host_dev_init(int *ncu) {
    const int rc = dev_init();
    if (ncu) *ncu = g_ncu;
    return rc;
}
int host_check_model(const rfba3b_model *m) { return check_model(m); }
std::mutex &host_mutex() { return g_mu; }
}  // namespace rfba3b

// ---------------------------------------------------------------------------
// batched plan

struct rfba3b_batch {
    rfba3b_model model;
    std::vector<rfba3b_block> ublocks;
    std::vector<DevBlock> blocks;
    // Chunks of blocks whose residual planes fit the plane buffer together.
    // Planes, items and correlation tasks are indexed chunk-locally on the
    // device (the chunk's first entry is passed as the array base).
    struct Chunk {
        int task0 = 0, ntask = 0, item0 = 0, nitem = 0, plane0 = 0, nplane = 0, corr0 = 0, ncorr = 0;
        // per (block, q): residual tasks, correlation tasks (global index),
        // items (chunk-local index)
        std::vector<int> bq_task0, bq_ntask, bq_corr0, bq_ncorr, bq_item0, bq_nitem, bq_block;
    };
    std::vector<Chunk> chunks;
    bool grad = true;
    int hmm_form = 0;  // rfba3b_batch_set_hmm_form
    int Tmax = 1, naxmax = 1, ckslots = 1;
    size_t lds_resid = 0, lds_corr = 0;
    DevModel *d_model = nullptr;
    DevBlock *d_blocks = nullptr;
    PlaneTask *d_tasks = nullptr;
    PlaneRef *d_planes = nullptr;
    CorrTask *d_corr = nullptr;
    Item *d_items = nullptr;
    float *d_R = nullptr;
    double *d_S0 = nullptr, *d_em = nullptr, *d_ck = nullptr;
    size_t R_floats = 0, S0_doubles = 0, em_doubles = 0, ck_doubles = 0;
    std::vector<PlaneTask> tasks;
    std::vector<PlaneRef> planes;
    std::vector<CorrTask> corr;
    std::vector<Item> items;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

static void batch_free(rfba3b_batch *b) {
    if (!b) return;
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    hipFree(b->d_model);
    hipFree(b->d_blocks);
    hipFree(b->d_tasks);
    hipFree(b->d_planes);
    hipFree(b->d_corr);
    hipFree(b->d_items);
    hipFree(b->d_R);
    hipFree(b->d_S0);
    hipFree(b->d_em);
    hipFree(b->d_ck);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    delete b;
}

typedef void (*CorrFn)(CorrParams);
typedef void (*HmmFn)(HmmParams);
static CorrFn corr_fn(bool grad) { return grad ? b3_corr_mfma<true, false, 0> : b3_corr_mfma<false, false, 0>; }
// the resident functor's form: region mask and pixel splits; for one or two
// candidates (at most 12 basis columns) matrix work on the first 16-column
// tile only
static CorrFn corr_fn_functor(int m) { return m <= 2 ? b3_corr_mfma<true, true, 1> : b3_corr_mfma<true, true, 0>; }

// The HMMs of a launch: one wave per item (b3_hmm_wave, the frames of an
// item spread over the lanes) when every block's T fits the wave's
// kWaveTmax frames, else one lane per item (b3_hmm, a serial chain of T
// frames per item).  The choice depends on T only, so that a batched
// launch and one launch per block or parameter set give the same bits.
// Measured on MI355X (29^3 block, T = 100): a functor call of 10 items
// takes 0.39 ms with the lane form and 0.15 ms with the wave form; one
// launch of 10240 items costs the same in both forms.
static int launch_hmm(int mode, const HmmParams &H, int Tmax, hipStream_t st, bool lane_form = false) {
    if (H.n <= 0) return 0;
    if (Tmax <= kWaveTmax && !lane_form) {
        const HmmFn fn = mode == MODE_GRAD    ? b3_hmm_wave<MODE_GRAD>
                         : mode == MODE_DELTA ? b3_hmm_wave<MODE_DELTA>
                                              : b3_hmm_wave<MODE_LOGL>;
        hipLaunchKernelGGL(fn, dim3((H.n + 3) / 4), dim3(256), 0, st, H);
    } else {
        const HmmFn fn = mode == MODE_GRAD ? b3_hmm<MODE_GRAD> : mode == MODE_DELTA ? b3_hmm<MODE_DELTA> : b3_hmm<MODE_LOGL>;
        hipLaunchKernelGGL(fn, dim3((H.n + 255) / 256), dim3(256), 0, st, H);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

static size_t corr_lds(bool grad, int naxmax) {
    // Q/H [8][6] | per-wave Q/H [4][8][6] | residual chunk [128][36] |
    // basis chunk [32][16-column tiles] | axis tables [3][nax][8]
    const size_t bmat = (size_t)kCorrPix * (((grad ? 48 : 8) + 15) / 16 * 16);
    return sizeof(double) * kCorrSpots * 6 * 5 +
           sizeof(float) * ((size_t)kCorrFrames * kCorrRow + bmat + (size_t)3 * kCorrSpots * naxmax);
}

extern "C" int rfba3b_batch_create(const rfba3b_block *blocks, int n, const rfba3b_model *model,
                                   rfba3b_batch **plan) {
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if (!plan) return fail(-1, "plan pointer is NULL");
    *plan = nullptr;
    if (!blocks || n <= 0) return fail(-1, "no blocks");
    if ((rc = check_model(model)) || (rc = dev_init())) return rc;
    rfba3b_batch *b = new rfba3b_batch();
    b->model = *model;
    bool grad = true;
    for (int i = 0; i < n; i++) {
        const rfba3b_block &u = blocks[i];
        if (u.dims[0] <= 0 || u.dims[1] <= 0 || u.dims[2] <= 0 || u.T <= 0 || u.K <= 0 || u.Q <= 0 || u.S <= 0 ||
            !u.d_data || !u.d_params || !u.d_states || !u.d_logL) {
            batch_free(b);
            return fail(-1, "block %d: invalid dimensions or buffers", i);
        }
        if (u.dims[0] + u.dims[1] + u.dims[2] > 1024 || (long long)u.dims[0] * u.dims[1] * u.dims[2] > (1 << 24) ||
            u.T > 4096) {
            batch_free(b);
            return fail(-1, "block %d: box or frame count too large", i);
        }
        grad = grad && u.d_grad;
        DevBlock d;
        memset(&d, 0, sizeof d);
        d.X = u.dims[0]; d.Y = u.dims[1]; d.Z = u.dims[2];
        d.T = u.T; d.K = u.K; d.Q = u.Q; d.S = u.S;
        d.b0 = u.b0;
        d.npix = u.dims[0] * u.dims[1] * u.dims[2];
        d.data = u.d_data; d.params = u.d_params; d.states = u.d_states;
        d.logL = u.d_logL; d.grad = u.d_grad;
        b->blocks.push_back(d);
        b->ublocks.push_back(u);
        b->Tmax = std::max(b->Tmax, u.T);
        b->naxmax = std::max(b->naxmax, u.dims[0] + u.dims[1] + u.dims[2]);
        // residual kernel tables: xy factors (padded to 4), z factors (Z
        // rounded up to 8 per spot), amp, on-state masks (two words per 64
        // spots)
        const size_t lr = sizeof(float) * ((((size_t)u.K * (u.dims[0] + u.dims[1]) + 3) & ~(size_t)3) +
                                           (size_t)u.K * ((u.dims[2] + 7) & ~7) + (size_t)u.K +
                                           2 * (((size_t)u.K + 63) / 64));
        b->lds_resid = std::max(b->lds_resid, lr);
        const int ck = ckpt_of(model, u.T);
        b->ckslots = std::max(b->ckslots, 6 * ((u.T + ck - 1) / ck) + 4 * ck);
    }
    if (b->lds_resid > 64 * 1024) {
        batch_free(b);
        return fail(-1, "too many spots per block for the residual kernel");
    }
    b->grad = grad;
    b->lds_corr = corr_lds(grad, b->naxmax);
    if (b->lds_corr > 160 * 1024) {
        batch_free(b);
        return fail(-1, "block edges too long for the correlation kernel");
    }

    // chunks: residual planes of all (q, s) of a block stay together; blocks
    // are added to a chunk while the planes fit the budget.
    const char *bud = getenv("RFBA3B_PLANE_BUDGET_MB");
    const size_t budget = (size_t)(bud ? atol(bud) : 16384) << 20;
    size_t cur = 0;
    rfba3b_batch::Chunk ch;
    size_t maxR = 0, maxPlanes = 0, maxItems = 0;
    auto close_chunk = [&]() {
        ch.ntask = (int)b->tasks.size() - ch.task0;
        ch.nitem = (int)b->items.size() - ch.item0;
        ch.nplane = (int)b->planes.size() - ch.plane0;
        ch.ncorr = (int)b->corr.size() - ch.corr0;
        maxR = std::max(maxR, cur);
        maxPlanes = std::max(maxPlanes, (size_t)ch.nplane);
        maxItems = std::max(maxItems, (size_t)ch.nitem);
        b->chunks.push_back(ch);
        ch = rfba3b_batch::Chunk();
        ch.task0 = (int)b->tasks.size();
        ch.item0 = (int)b->items.size();
        ch.plane0 = (int)b->planes.size();
        ch.corr0 = (int)b->corr.size();
        cur = 0;
    };
    for (int bi = 0; bi < n; bi++) {
        const DevBlock &d = b->blocks[bi];
        const size_t P = (size_t)d.X * d.Y * d.Z;
        const size_t bytes = (size_t)d.Q * d.S * d.T * P * sizeof(float);
        if (cur > 0 && cur + bytes > budget) close_chunk();
        for (int q = 0; q < d.Q; q++) {
            ch.bq_block.push_back(bi);
            ch.bq_task0.push_back((int)b->tasks.size());
            ch.bq_corr0.push_back((int)b->corr.size());
            ch.bq_item0.push_back((int)b->items.size() - ch.item0);
            for (int s = 0; s < d.S; s++) {
                const long long off = (long long)(cur / sizeof(float));
                const int plane = (int)b->planes.size() - ch.plane0;
                const int item0 = (int)b->items.size() - ch.item0;
                b->planes.push_back({bi, q, s, 0, off});
                for (int t = 0; t < d.T; t++) b->tasks.push_back({bi, q, s, t, off, plane, 0});
                for (int k = 0; k < d.K; k++) b->items.push_back({bi, q, k, s, off});
                for (int k0 = 0; k0 < d.K; k0 += kCorrSpots)
                    b->corr.push_back({plane, k0, std::min(kCorrSpots, d.K - k0), item0 + k0});
                cur += (size_t)d.T * P * sizeof(float);
            }
            ch.bq_ntask.push_back((int)b->tasks.size() - ch.bq_task0.back());
            ch.bq_ncorr.push_back((int)b->corr.size() - ch.bq_corr0.back());
            ch.bq_nitem.push_back((int)b->items.size() - ch.item0 - ch.bq_item0.back());
        }
    }
    close_chunk();
    b->R_floats = maxR / sizeof(float);
    b->S0_doubles = maxPlanes * (size_t)b->Tmax;
    b->em_doubles = maxItems * (size_t)b->Tmax * 8;
    b->ck_doubles = maxItems * (size_t)b->ckslots;

    if (b->lds_corr > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)corr_fn(grad), hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)b->lds_corr);
        if (e != hipSuccess) {
            batch_free(b);
            return fail(-3, "correlation kernel LDS: %s", hipGetErrorString(e));
        }
    }
    hipError_t e = hipSuccess;
    DevModel dm = dev_model(model);
    auto al = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 16));
    };
    al((void **)&b->d_model, sizeof dm);
    al((void **)&b->d_blocks, b->blocks.size() * sizeof(DevBlock));
    al((void **)&b->d_tasks, b->tasks.size() * sizeof(PlaneTask));
    al((void **)&b->d_planes, b->planes.size() * sizeof(PlaneRef));
    al((void **)&b->d_corr, b->corr.size() * sizeof(CorrTask));
    al((void **)&b->d_items, b->items.size() * sizeof(Item));
    al((void **)&b->d_R, b->R_floats * sizeof(float));
    al((void **)&b->d_S0, b->S0_doubles * sizeof(double));
    al((void **)&b->d_em, b->em_doubles * sizeof(double));
    al((void **)&b->d_ck, b->ck_doubles * sizeof(double));
    auto up = [&](void *d, const void *h, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpy(d, h, bytes, hipMemcpyHostToDevice);
    };
    up(b->d_model, &dm, sizeof dm);
    up(b->d_blocks, b->blocks.data(), b->blocks.size() * sizeof(DevBlock));
    up(b->d_tasks, b->tasks.data(), b->tasks.size() * sizeof(PlaneTask));
    up(b->d_planes, b->planes.data(), b->planes.size() * sizeof(PlaneRef));
    up(b->d_corr, b->corr.data(), b->corr.size() * sizeof(CorrTask));
    up(b->d_items, b->items.data(), b->items.size() * sizeof(Item));
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e != hipSuccess) {
        int code = e == hipErrorOutOfMemory ? -2 : -3;
        batch_free(b);
        return fail(code, "3B plan setup: %s", hipGetErrorString(e));
    }
    *plan = b;
    return 0;
}

// One chunk's work, or the part of it that belongs to one (block, q):
// residual planes (+ per-frame sums of squares), correlations, HMMs.
static int enqueue_range(rfba3b_batch *b, hipStream_t st, const rfba3b_batch::Chunk &c, int task0, int ntask,
                         int corr0, int ncorr, int ilo, int nitem) {
    if (ntask > 0) {
        const int gr = std::min(ntask, g_ncu * 8);
        hipLaunchKernelGGL(b3_resid, dim3(gr), dim3(256), b->lds_resid, st, (const DevBlock *)b->d_blocks,
                           (const PlaneTask *)(b->d_tasks + task0), ntask, b->d_R, b->d_S0, b->Tmax,
                           (const DevModel *)b->d_model);
        HIP_TRY(hipGetLastError());
    }
    if (ncorr > 0) {
        CorrParams C;
        memset(&C, 0, sizeof C);
        C.blocks = b->d_blocks;
        C.planes = b->d_planes + c.plane0;
        C.tasks = b->d_corr + corr0;
        C.ntasks = ncorr;
        C.naxmax = b->naxmax;
        C.Tmax = b->Tmax;
        C.NI = c.nitem;
        C.R = b->d_R;
        C.S0 = b->d_S0;
        C.model = b->d_model;
        C.em = b->d_em;
        hipLaunchKernelGGL(corr_fn(b->grad), dim3((ncorr + 7) / 8 * 8), dim3(256), b->lds_corr, st, C);
        HIP_TRY(hipGetLastError());
    }
    if (nitem > 0) {
        HmmParams H;
        memset(&H, 0, sizeof H);
        H.blocks = b->d_blocks;
        H.items = b->d_items + c.item0;
        H.lo = ilo;
        H.n = nitem;
        H.NI = c.nitem;
        H.ckpt = b->model.checkpoint;
        H.ckslots = b->ckslots;
        H.em = b->d_em;
        H.ck = b->d_ck;
        H.model = b->d_model;
        const int rc = launch_hmm(b->grad ? MODE_GRAD : MODE_LOGL, H, b->Tmax, st, b->hmm_form == 1);
        if (rc) return rc;
    }
    return 0;
}

static int enqueue_all(rfba3b_batch *b, hipStream_t st) {
    int rc;
    for (auto &c : b->chunks)
        if ((rc = enqueue_range(b, st, c, c.task0, c.ntask, c.corr0, c.ncorr, 0, c.nitem))) return rc;
    return 0;
}

extern "C" int rfba3b_batch_run(rfba3b_batch *b, void *stream, int use_graph, float *kernel_ms) {
    if (!b) return fail(-1, "plan is NULL");
    std::lock_guard<std::mutex> lk(g_mu);
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if (use_graph) {
        if (!b->exec) {
            hipStream_t cs;
            HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            rc = enqueue_all(b, cs);
            hipGraph_t gr = nullptr;
            hipError_t e = hipStreamEndCapture(cs, &gr);
            hipStreamDestroy(cs);
            if (rc) return rc;
            if (e != hipSuccess) return fail(-3, "graph capture: %s", hipGetErrorString(e));
            b->graph = gr;
            HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        HIP_TRY(hipEventRecord(b->ev0, st));
        HIP_TRY(hipGraphLaunch(b->exec, st));
        HIP_TRY(hipEventRecord(b->ev1, st));
    } else {
        HIP_TRY(hipEventRecord(b->ev0, st));
        if ((rc = enqueue_all(b, st))) return rc;
        HIP_TRY(hipEventRecord(b->ev1, st));
    }
    if (kernel_ms) {
        HIP_TRY(hipEventSynchronize(b->ev1));
        HIP_TRY(hipEventElapsedTime(kernel_ms, b->ev0, b->ev1));
    }
    return 0;
}

extern "C" int rfba3b_batch_run_per_block(rfba3b_batch *b, void *stream, int per_param_set, float *kernel_ms) {
    if (!b) return fail(-1, "plan is NULL");
    std::lock_guard<std::mutex> lk(g_mu);
    hipStream_t st = (hipStream_t)stream;
    int rc;
    HIP_TRY(hipEventRecord(b->ev0, st));
    for (auto &c : b->chunks)
        for (size_t i = 0; i < c.bq_task0.size();) {
            // ranges of one (block, q), or of all q of one block (consecutive)
            size_t j = i + 1;
            if (!per_param_set)
                while (j < c.bq_task0.size() && c.bq_block[j] == c.bq_block[i]) j++;
            const int nt = c.bq_task0[j - 1] + c.bq_ntask[j - 1] - c.bq_task0[i];
            const int nc = c.bq_corr0[j - 1] + c.bq_ncorr[j - 1] - c.bq_corr0[i];
            const int ni = c.bq_item0[j - 1] + c.bq_nitem[j - 1] - c.bq_item0[i];
            if ((rc = enqueue_range(b, st, c, c.bq_task0[i], nt, c.bq_corr0[i], nc, c.bq_item0[i], ni))) return rc;
            i = j;
        }
    HIP_TRY(hipEventRecord(b->ev1, st));
    if (kernel_ms) {
        HIP_TRY(hipEventSynchronize(b->ev1));
        HIP_TRY(hipEventElapsedTime(kernel_ms, b->ev0, b->ev1));
    }
    return 0;
}

extern "C" int rfba3b_batch_chunks(const rfba3b_batch *b) {
    if (!b) return fail(-1, "plan is NULL");
    return (int)b->chunks.size();
}

extern "C" int rfba3b_batch_set_hmm_form(rfba3b_batch *b, int form) {
    if (!b) return fail(-1, "plan is NULL");
    if (form != 0 && form != 1) return fail(-1, "HMM form %d: 0 (by series length) or 1 (lane form)", form);
    std::lock_guard<std::mutex> lk(g_mu);
    const int prev = b->hmm_form;
    if (form != prev) {
        // a captured graph holds the other form's launches
        if (b->exec) hipGraphExecDestroy(b->exec);
        if (b->graph) hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        b->hmm_form = form;
    }
    return prev;
}

extern "C" void rfba3b_batch_destroy(rfba3b_batch *b) {
    std::lock_guard<std::mutex> lk(g_mu);
    batch_free(b);
}

// ---------------------------------------------------------------------------
// functor and forward messages over explicit backgrounds

namespace {

inline double ln0(double x) { return x == 0 ? -1e100 : std::log(x); }
inline double dlln(double x, double mu, double sigma) { return -(1 + (ln0(x) - mu) / (sigma * sigma)) / x; }

// Explicit-background evaluation: candidates phi[n][5] of one spot over S
// backgrounds.  Returns logL [n][S] and grad [n][S][5] (mode GRAD) or the
// forward messages [n*S][T][3] (mode DELTA).
int eval_explicit(const int *xyz, int P, int T, const float *data, const float *bg, int S, const double *phi, int n,
                  const rfba3b_model *m, int mode, std::vector<double> &logL, std::vector<double> &grad,
                  std::vector<double> *delta) {
    int rc;
    if ((rc = check_model(m)) || (rc = dev_init())) return rc;
    if (!xyz || !data || !bg || !phi || P <= 0 || T <= 0 || S <= 0 || n <= 0) return fail(-1, "invalid arguments");
    int lo[3] = {xyz[0], xyz[1], xyz[2]}, hi[3] = {xyz[0], xyz[1], xyz[2]};
    for (int p = 0; p < P; p++)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], xyz[3 * p + a]);
            hi[a] = std::max(hi[a], xyz[3 * p + a]);
        }
    const int X = hi[0] - lo[0] + 1, Y = hi[1] - lo[1] + 1, Z = hi[2] - lo[2] + 1;
    if ((long long)X * Y > 4096 || Z > 1024 || T > 4096) return fail(-1, "pixel region too large");
    const size_t Pd = (size_t)X * Y * Z;
    std::vector<uint8_t> mask(Pd, 0);
    std::vector<int> map(P);
    for (int p = 0; p < P; p++) {
        const size_t d = ((size_t)(xyz[3 * p + 2] - lo[2]) * Y + (xyz[3 * p + 1] - lo[1])) * X + (xyz[3 * p] - lo[0]);
        if (mask[d]) return fail(-1, "duplicate pixel in the region");
        mask[d] = 1;
        map[p] = (int)d;
    }
    std::vector<float> R((size_t)S * T * Pd, 0.0f);
    for (int s = 0; s < S; s++)
        for (int t = 0; t < T; t++)
            for (int p = 0; p < P; p++)
                R[((size_t)s * T + t) * Pd + map[p]] = data[(size_t)t * P + p] - bg[((size_t)s * T + t) * P + p];
    DevBlock d;
    memset(&d, 0, sizeof d);
    d.X = X; d.Y = Y; d.Z = Z; d.T = T; d.K = n; d.Q = 1; d.S = S;
    d.ox = lo[0]; d.oy = lo[1]; d.oz = lo[2];
    d.npix = P;
    std::vector<Item> items;
    for (int k = 0; k < n; k++)
        for (int s = 0; s < S; s++) items.push_back({0, 0, k, s, (long long)s * T * (long long)Pd});
    const int ckmax = ckpt_of(m, T);
    const size_t lds = eval_lds(T, X * Y, Z, ckmax);
    EvalFn fn = eval_fn(mode);
    int grid = 0;
    if ((rc = eval_grid(fn, lds, (int)items.size(), grid))) return rc;

    float *dR = nullptr;
    uint8_t *dmask = nullptr;
    double *dphi = nullptr, *dL = nullptr, *dG = nullptr, *dD = nullptr;
    DevModel *dm = nullptr;
    DevBlock *db = nullptr;
    Item *di = nullptr;
    int *dc = nullptr;
    const DevModel hm = dev_model(m);
    auto cleanup = [&] {
        hipFree(dR); hipFree(dmask); hipFree(dphi); hipFree(dL); hipFree(dG); hipFree(dD);
        hipFree(dm); hipFree(db); hipFree(di); hipFree(dc);
    };
    hipError_t e = hipSuccess;
    const size_t nitem = items.size();
    if (e == hipSuccess) e = hipMalloc(&dR, R.size() * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dmask, Pd);
    if (e == hipSuccess) e = hipMalloc(&dphi, (size_t)n * 5 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dL, nitem * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dG, nitem * 5 * sizeof(double));
    if (e == hipSuccess && mode == MODE_DELTA) e = hipMalloc(&dD, nitem * 3 * (size_t)T * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dm, sizeof(DevModel));
    if (e == hipSuccess) e = hipMalloc(&db, sizeof(DevBlock));
    if (e == hipSuccess) e = hipMalloc(&di, nitem * sizeof(Item));
    if (e == hipSuccess) e = hipMalloc(&dc, sizeof(int));
    d.params = dphi;
    d.mask = dmask;
    d.logL = dL;
    d.grad = dG;
    if (e == hipSuccess) e = hipMemcpy(dR, R.data(), R.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dmask, mask.data(), Pd, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dphi, phi, (size_t)n * 5 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dm, &hm, sizeof hm, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, &d, sizeof d, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(di, items.data(), nitem * sizeof(Item), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        cleanup();
        return fail(e == hipErrorOutOfMemory ? -2 : -3, "3B setup: %s", hipGetErrorString(e));
    }
    EvalParams E;
    memset(&E, 0, sizeof E);
    E.blocks = db; E.items = di; E.nitems = (int)nitem; E.counter = dc; E.R = dR; E.model = dm;
    E.rmode = 1; E.ckpt = m->checkpoint; E.Tmax = T; E.XYmax = X * Y; E.Zmax = Z; E.delta_out = dD;
    hipLaunchKernelGGL(b3_reset, dim3(1), dim3(64), 0, nullptr, dc);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(256), lds, nullptr, E);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    logL.resize(nitem);
    grad.resize(nitem * 5);
    if (e == hipSuccess) e = hipMemcpy(logL.data(), dL, nitem * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && mode == MODE_GRAD)
        e = hipMemcpy(grad.data(), dG, nitem * 5 * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && mode == MODE_DELTA) {
        delta->resize(nitem * 3 * (size_t)T);
        e = hipMemcpy(delta->data(), dD, delta->size() * sizeof(double), hipMemcpyDeviceToHost);
    }
    cleanup();
    if (e != hipSuccess) return fail(-3, "3B evaluation: %s", hipGetErrorString(e));
    return 0;
}

}  // namespace

extern "C" int rfba3b_neg_grad_sampled(const int *xyz, int P, int T, const float *data, const float *bg, int S,
                                       const double *phi, int n, const rfba3b_model *model, double *grad,
                                       double *logL) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!grad || !phi || n <= 0) return fail(-1, "invalid arguments");
    // multispot5.cc:618-619: invalid parameters give a NaN vector.
    std::vector<double> good;
    std::vector<int> idx;
    for (int i = 0; i < n; i++)
        if (phi[5 * i] > 0 && phi[5 * i + 1] > 0) {
            idx.push_back(i);
            good.insert(good.end(), phi + 5 * i, phi + 5 * i + 5);
        }
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (int i = 0; i < n; i++) {
        for (int c = 0; c < 5; c++) grad[5 * i + c] = nan;
        if (logL)
            for (int s = 0; s < S; s++) logL[(size_t)i * S + s] = nan;
    }
    if (idx.empty()) return 0;
    std::vector<double> L, G;
    int rc = eval_explicit(xyz, P, T, data, bg, S, good.data(), (int)idx.size(), model, MODE_GRAD, L, G, nullptr);
    if (rc) return rc;
    for (size_t c = 0; c < idx.size(); c++) {
        const int i = idx[c];
        double sum[5] = {0, 0, 0, 0, 0};
        for (int s = 0; s < S; s++) {
            for (int k = 0; k < 5; k++) sum[k] += G[(c * S + s) * 5 + k];
            if (logL) logL[(size_t)i * S + s] = L[c * S + s];
        }
        const double *p = phi + 5 * i;
        const double prior[5] = {dlln(p[0], model->mu_brightness, model->sigma_brightness),
                                 dlln(p[1], model->mu_blur, model->sigma_blur), 0, 0, 0};
        for (int k = 0; k < 5; k++) grad[5 * i + k] = -(sum[k] / S + prior[k]);
    }
    return 0;
}

extern "C" int rfba3b_forward_messages(const int *xyz, int P, int T, const float *data, const float *bg,
                                       const double *phi, const rfba3b_model *model, double *delta, double *logL) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!delta || !phi) return fail(-1, "invalid arguments");
    if (!(phi[0] > 0 && phi[1] > 0)) return fail(-1, "invalid spot parameters");
    std::vector<double> L, G, D;
    int rc = eval_explicit(xyz, P, T, data, bg, 1, phi, 1, model, MODE_DELTA, L, G, &D);
    if (rc) return rc;
    memcpy(delta, D.data(), 3 * (size_t)T * sizeof(double));
    if (logL) *logL = L[0];
    return 0;
}

// ---------------------------------------------------------------------------
// Resident CG functor: SampledBackgroundData (multispot5.cc:538-575) kept on
// the device, evaluated by the batched correlation and HMM kernels.
//
// Per background sample s the residual data - bg_s is formed at creation in
// FP64 and rounded once to FP32 (one plane [T][box], zero outside the
// region), with its per-frame sum of squares S0_s(t) in FP64.  A candidate
// spot is then an item (spot k, sample s) of one virtual block whose states
// are all "off": the correlation kernel's off-state sums are exactly the
// terms of spot_with_background.hh:127-137 with e = R - mu
// (Soff = S0, D = 2C - Q, diff = G - H), and the HMM kernel gives logL and
// its gradient.  An evaluation uploads the candidates and the task lists
// into buffers sized at creation, launches two kernels and reads back
// [n][S] values: it allocates nothing.

struct rfba3b_functor {
    rfba3b_model model;
    int P = 0, T = 0, S = 0, nmax = 0, X = 0, Y = 0, Z = 0, ckslots = 0;
    size_t lds_corr = 0;
    DevModel *d_model = nullptr;
    DevBlock *d_block = nullptr;
    PlaneRef *d_planes = nullptr;
    CorrTask *d_corr = nullptr;  // in d_stage
    Item *d_items = nullptr;     // in d_stage
    float *d_R = nullptr;
    uint8_t *d_states = nullptr, *d_mask = nullptr;
    double *d_S0 = nullptr, *d_em = nullptr, *d_ck = nullptr;
    // staging, pinned on the host and mirrored in one device allocation:
    // params [nmax][5] | correlation tasks | items | logL [nmax][S] |
    // grad [nmax][S][5], so that a call uploads its inputs in one copy and
    // downloads its outputs in one copy
    unsigned char *h_stage = nullptr, *d_stage = nullptr;
    size_t o_logL = 0, o_grad = 0, o_corr = 0, o_items = 0;
    double *d_params = nullptr, *d_logL = nullptr, *d_grad = nullptr;  // in d_stage
    std::mutex mu;
    // pixel-split correlation sums (few candidates: each plane's pixels are
    // spread over many workgroups)
    double *d_part = nullptr, *d_part_qh = nullptr;
    size_t part_doubles = 0, partqh_doubles = 0;
    // FFBS Hessian buffers, allocated by the first rfba3b_functor_hessian_ffbs
    // and grown with bs_iterations: queue counter, spot table and its
    // d d^T partial sums, per-frame parts, forward messages, draws, paths
    int *d_counter = nullptr;
    float *d_basis = nullptr;
    double *d_fixed = nullptr, *d_parts = nullptr, *d_delta = nullptr, *d_unif = nullptr;
    uint8_t *d_paths = nullptr;
    size_t paths_cap = 0;  // chains (S x bs) the draw and path buffers hold
};

static void functor_free(rfba3b_functor *f) {
    if (!f) return;
    hipFree(f->d_part); hipFree(f->d_part_qh);
    hipFree(f->d_counter); hipFree(f->d_basis); hipFree(f->d_fixed); hipFree(f->d_parts); hipFree(f->d_delta);
    hipFree(f->d_unif); hipFree(f->d_paths);
    hipFree(f->d_model); hipFree(f->d_block); hipFree(f->d_planes); hipFree(f->d_stage);
    hipFree(f->d_R); hipFree(f->d_states); hipFree(f->d_mask); hipFree(f->d_S0); hipFree(f->d_em);
    hipFree(f->d_ck);
    if (f->h_stage) hipHostFree(f->h_stage);
    delete f;
}

// Pixel splits of the functor's correlation for m candidates: enough
// (plane, spot group, pixel range) tasks for about three workgroups per CU,
// each range at least 8 chunks of 32 pixels, and the partial sums within the
// handle's buffer; 1 = whole planes.
static void functor_split(const rfba3b_functor *f, int m, int &nsplit, int &pstep) {
    const long long Pd = (long long)f->X * f->Y * f->Z;
    const long long base = (long long)f->S * ((m + kCorrSpots - 1) / kCorrSpots);
    const long long NI = (long long)f->S * m;
    long long ns = std::min<long long>((3LL * g_ncu + base - 1) / base, (Pd + 8 * kCorrPix - 1) / (8 * kCorrPix));
    if (f->part_doubles) ns = std::min<long long>(ns, (long long)(f->part_doubles / ((size_t)NI * f->T * 6)));
    if (f->partqh_doubles) ns = std::min<long long>(ns, (long long)(f->partqh_doubles / ((size_t)NI * 6)));
    if (ns < 2) {
        nsplit = 1;
        pstep = (int)Pd;
        return;
    }
    const long long per = (Pd + ns - 1) / ns;
    pstep = (int)((per + kCorrPix - 1) / kCorrPix * kCorrPix);
    nsplit = (int)((Pd + pstep - 1) / pstep);
}

extern "C" int rfba3b_functor_create(const int *xyz, int P, int T, const double *data, const double *bg, int S,
                                     const rfba3b_model *model, int max_candidates, rfba3b_functor **out) {
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if (!out) return fail(-1, "handle pointer is NULL");
    *out = nullptr;
    if ((rc = check_model(model)) || (rc = dev_init())) return rc;
    if (!xyz || !data || !bg || P <= 0 || T <= 0 || S <= 0 || max_candidates <= 0)
        return fail(-1, "invalid arguments");
    if (T > 4096 || max_candidates > 4096 || S > 4096) return fail(-1, "T, S and max_candidates must be <= 4096");
    int lo[3] = {xyz[0], xyz[1], xyz[2]}, hi[3] = {xyz[0], xyz[1], xyz[2]};
    for (int p = 0; p < P; p++)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], xyz[3 * p + a]);
            hi[a] = std::max(hi[a], xyz[3 * p + a]);
        }
    const long long Xl = (long long)hi[0] - lo[0] + 1, Yl = (long long)hi[1] - lo[1] + 1,
                    Zl = (long long)hi[2] - lo[2] + 1;
    if (Xl + Yl + Zl > 1024 || Xl * Yl * Zl > (1 << 24)) return fail(-1, "pixel region too large");
    const int X = (int)Xl, Y = (int)Yl, Z = (int)Zl;
    const size_t Pd = (size_t)X * Y * Z;
    std::vector<uint8_t> mask(Pd, 0);
    std::vector<size_t> map(P);
    for (int p = 0; p < P; p++) {
        const size_t d = ((size_t)(xyz[3 * p + 2] - lo[2]) * Y + (xyz[3 * p + 1] - lo[1])) * X + (xyz[3 * p] - lo[0]);
        if (mask[d]) return fail(-1, "duplicate pixel in the region");
        mask[d] = 1;
        map[p] = d;
    }
    const bool full = (size_t)P == Pd;
    // residual planes and their per-frame sums of squares
    std::vector<float> R((size_t)S * T * Pd, 0.0f);
    std::vector<double> S0((size_t)S * T, 0.0);
    for (int s = 0; s < S; s++)
        for (int t = 0; t < T; t++) {
            float *row = R.data() + ((size_t)s * T + t) * Pd;
            const double *dr = data + (size_t)t * P, *br = bg + ((size_t)s * T + t) * P;
            for (int p = 0; p < P; p++) row[map[p]] = (float)(dr[p] - br[p]);
            double acc = 0.0;
            for (size_t d = 0; d < Pd; d++) acc += (double)row[d] * (double)row[d];
            S0[(size_t)s * T + t] = acc;
        }

    rfba3b_functor *f = new rfba3b_functor();
    f->model = *model;
    f->P = P; f->T = T; f->S = S; f->nmax = max_candidates;
    f->X = X; f->Y = Y; f->Z = Z;
    const int ck = ckpt_of(model, T);
    f->ckslots = 6 * ((T + ck - 1) / ck) + 4 * ck;
    f->lds_corr = corr_lds(true, X + Y + Z);
    if (f->lds_corr > 160 * 1024) {
        functor_free(f);
        return fail(-1, "region edges too long for the correlation kernel");
    }
    const size_t NI = (size_t)S * max_candidates;
    const size_t ncorr = (size_t)S * ((max_candidates + kCorrSpots - 1) / kCorrSpots);
    f->o_corr = (size_t)max_candidates * 5 * sizeof(double);
    f->o_items = f->o_corr + ncorr * sizeof(CorrTask);
    f->o_logL = (f->o_items + NI * sizeof(Item) + 7) / 8 * 8;
    f->o_grad = f->o_logL + NI * sizeof(double);
    const size_t stage = f->o_grad + NI * 5 * sizeof(double);
    // the largest partial-sum buffer any candidate count needs, at most 256 MiB
    {
        size_t need = 0, needqh = 0;
        const size_t cap = (size_t)32 << 20;  // doubles
        for (int m = 1; m <= max_candidates; m++) {
            int ns, ps;
            functor_split(f, m, ns, ps);
            if (ns < 2) continue;
            need = std::max(need, (size_t)ns * S * m * T * 6);
            needqh = std::max(needqh, (size_t)ns * S * m * 6);
        }
        f->part_doubles = std::min(need, cap);
        f->partqh_doubles = std::min(needqh, cap);
    }

    hipError_t e = hipSuccess;
    auto al = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 16));
    };
    al((void **)&f->d_model, sizeof(DevModel));
    al((void **)&f->d_block, sizeof(DevBlock));
    al((void **)&f->d_planes, S * sizeof(PlaneRef));
    al((void **)&f->d_stage, stage);
    al((void **)&f->d_R, R.size() * sizeof(float));
    al((void **)&f->d_states, NI * (size_t)T);
    if (!full) al((void **)&f->d_mask, Pd);
    al((void **)&f->d_S0, S0.size() * sizeof(double));
    al((void **)&f->d_em, NI * (size_t)T * 8 * sizeof(double));
    al((void **)&f->d_ck, NI * (size_t)f->ckslots * sizeof(double));
    if (e == hipSuccess) {
        f->d_params = (double *)f->d_stage;
        f->d_corr = (CorrTask *)(f->d_stage + f->o_corr);
        f->d_items = (Item *)(f->d_stage + f->o_items);
        f->d_logL = (double *)(f->d_stage + f->o_logL);
        f->d_grad = (double *)(f->d_stage + f->o_grad);
    }
    if (f->part_doubles) {
        al((void **)&f->d_part, f->part_doubles * sizeof(double));
        al((void **)&f->d_part_qh, f->partqh_doubles * sizeof(double));
    }
    if (e == hipSuccess) e = hipHostMalloc((void **)&f->h_stage, stage, hipHostMallocDefault);
    DevBlock d;
    memset(&d, 0, sizeof d);
    d.X = X; d.Y = Y; d.Z = Z; d.T = T; d.K = max_candidates; d.Q = 1; d.S = S;
    d.ox = lo[0]; d.oy = lo[1]; d.oz = lo[2];
    d.npix = P;
    d.params = f->d_params;
    d.states = f->d_states;
    d.mask = f->d_mask;
    d.logL = f->d_logL;
    d.grad = f->d_grad;
    std::vector<PlaneRef> planes(S);
    for (int s = 0; s < S; s++) planes[s] = {0, 0, s, 0, (long long)s * T * (long long)Pd};
    const DevModel dm = dev_model(model);
    auto up = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    };
    up(f->d_model, &dm, sizeof dm);
    up(f->d_block, &d, sizeof d);
    up(f->d_planes, planes.data(), S * sizeof(PlaneRef));
    up(f->d_R, R.data(), R.size() * sizeof(float));
    up(f->d_S0, S0.data(), S0.size() * sizeof(double));
    if (!full) up(f->d_mask, mask.data(), Pd);
    if (e == hipSuccess) e = hipMemset(f->d_states, 1, NI * (size_t)T);  // every candidate off in every frame
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess && f->lds_corr > 64 * 1024)
        for (int m : {1, 3})
            if (e == hipSuccess)
                e = hipFuncSetAttribute((const void *)corr_fn_functor(m), hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)f->lds_corr);
    if (e != hipSuccess) {
        const int code = e == hipErrorOutOfMemory ? -2 : -3;
        functor_free(f);
        return fail(code, "functor setup: %s", hipGetErrorString(e));
    }
    *out = f;
    return 0;
}

// The per-frame sums em of the m candidates staged at the start of h_stage
// over the handle's S background samples (item s m + k: candidate k over
// background s): correlation, then the reduction of the pixel splits.
// Called with g_mu held.
static int functor_sums(rfba3b_functor *f, int m, hipStream_t st) {
    const int S = f->S, T = f->T;
    CorrTask *ct = (CorrTask *)(f->h_stage + f->o_corr);
    Item *it = (Item *)(f->h_stage + f->o_items);
    const long long plane = (long long)T * f->X * f->Y * f->Z;
    int nc = 0;
    for (int s = 0; s < S; s++) {
        for (int k0 = 0; k0 < m; k0 += kCorrSpots) ct[nc++] = {s, k0, std::min(kCorrSpots, m - k0), s * m + k0};
        for (int k = 0; k < m; k++) it[s * m + k] = {0, 0, k, s, s * plane};
    }
    const int NI = S * m;
    // params, tasks and items in one copy (unused slots between them included)
    HIP_TRY(hipMemcpyAsync(f->d_stage, f->h_stage, f->o_items + NI * sizeof(Item), hipMemcpyHostToDevice, st));
    CorrParams C;
    memset(&C, 0, sizeof C);
    C.blocks = f->d_block;
    C.planes = f->d_planes;
    C.tasks = f->d_corr;
    C.ntasks = nc;
    C.naxmax = f->X + f->Y + f->Z;
    C.Tmax = T;
    C.NI = NI;
    C.R = f->d_R;
    C.S0 = f->d_S0;
    C.model = f->d_model;
    C.em = f->d_em;
    int nsplit = 1, pstep = 0;
    functor_split(f, m, nsplit, pstep);
    if (nsplit > 1) {
        C.nsplit = nsplit;
        C.pstep = pstep;
        C.part = f->d_part;
        C.part_qh = f->d_part_qh;
    }
    hipLaunchKernelGGL(corr_fn_functor(m), dim3((nc * nsplit + 7) / 8 * 8), dim3(256), f->lds_corr, st, C);
    HIP_TRY(hipGetLastError());
    if (nsplit > 1) {
        CorrReduceParams Rd;
        Rd.block = f->d_block;
        Rd.items = f->d_items;
        Rd.NI = NI;
        Rd.T = T;
        Rd.nsplit = nsplit;
        Rd.Tmax = T;
        Rd.part = f->d_part;
        Rd.part_qh = f->d_part_qh;
        Rd.S0 = f->d_S0;
        Rd.em = f->d_em;
        hipLaunchKernelGGL(b3_corr_reduce, dim3((NI * T * 6 * 8 + 255) / 256), dim3(256), 0, st, Rd);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

// The HMM launch over the NI items of functor_sums.
static HmmParams functor_hmm(const rfba3b_functor *f, int NI) {
    HmmParams H;
    memset(&H, 0, sizeof H);
    H.blocks = f->d_block;
    H.items = f->d_items;
    H.lo = 0;
    H.n = NI;
    H.NI = NI;
    H.ckpt = f->model.checkpoint;
    H.ckslots = f->ckslots;
    H.em = f->d_em;
    H.ck = f->d_ck;
    H.model = f->d_model;
    return H;
}

extern "C" int rfba3b_functor_eval(rfba3b_functor *f, const double *phi, int n, double *grad, double *logL,
                                   void *stream) {
    if (!f) return fail(-1, "handle is NULL");
    if (!phi || !grad || n <= 0) return fail(-1, "invalid arguments");
    if (n > f->nmax) return fail(-1, "%d candidates exceed the handle's maximum of %d", n, f->nmax);
    std::lock_guard<std::mutex> hl(f->mu);
    const int S = f->S, T = f->T;
    hipStream_t st = (hipStream_t)stream;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    // multispot5.cc:618-619: invalid parameters give a NaN vector
    double *sp = (double *)f->h_stage;
    int idx[4096];
    int m = 0;
    for (int i = 0; i < n; i++) {
        for (int c = 0; c < 5; c++) grad[5 * i + c] = nan;
        if (logL)
            for (int s = 0; s < S; s++) logL[(size_t)i * S + s] = nan;
        if (phi[5 * i] > 0 && phi[5 * i + 1] > 0) {
            for (int c = 0; c < 5; c++) sp[5 * m + c] = phi[5 * i + c];
            idx[m++] = i;
        }
    }
    if (m == 0) return 0;
    const int NI = S * m;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        int rc = functor_sums(f, m, st);
        if (!rc) rc = launch_hmm(MODE_GRAD, functor_hmm(f, NI), T, st);
        if (rc) return rc;
        // outputs [k][s] of the first m candidates (the block's K is nmax):
        // logL and grad in one copy
        HIP_TRY(hipMemcpyAsync(f->h_stage + f->o_logL, f->d_stage + f->o_logL,
                               f->o_grad - f->o_logL + (size_t)NI * 5 * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    const double *L = (const double *)(f->h_stage + f->o_logL);
    const double *G = (const double *)(f->h_stage + f->o_grad);
    for (int c = 0; c < m; c++) {
        const int i = idx[c];
        double sum[5] = {0, 0, 0, 0, 0};
        for (int s = 0; s < S; s++) {
            for (int k = 0; k < 5; k++) sum[k] += G[((size_t)c * S + s) * 5 + k];
            if (logL) logL[(size_t)i * S + s] = L[(size_t)c * S + s];
        }
        const double *p = phi + 5 * i;
        const double prior[5] = {dlln(p[0], f->model.mu_brightness, f->model.sigma_brightness),
                                 dlln(p[1], f->model.mu_blur, f->model.sigma_blur), 0, 0, 0};
        for (int k = 0; k < 5; k++) grad[5 * i + k] = -(sum[k] / S + prior[k]);
    }
    return 0;
}

extern "C" void rfba3b_functor_destroy(rfba3b_functor *f) {
    std::lock_guard<std::mutex> lk(g_mu);
    functor_free(f);
}

// ---------------------------------------------------------------------------
// forward filtering / backward sampling

static LogA log_a(const rfba3b_model *m) {
    LogA l;
    for (int i = 0; i < 9; i++) l.v[i] = ln0(m->A[i]);
    return l;
}

extern "C" int rfba3b_backward_sample_device(const double *d_delta, int n, int T, const rfba3b_model *model,
                                             const double *d_uniforms, uint8_t *d_states, void *stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    int rc;
    if ((rc = check_model(model)) || (rc = dev_init())) return rc;
    if (n < 0 || T <= 0 || (n > 0 && (!d_delta || !d_uniforms || !d_states))) return fail(-1, "invalid arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(b3_backward_sample, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, d_delta, n, 1,
                       T, log_a(model), d_uniforms, d_states);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int rfba3b_ffbs(const int *xyz, int P, int T, const float *data, const float *bg, const double *phi,
                           const rfba3b_model *model, const double *uniforms, uint8_t *states, double *delta,
                           double *logL) {
    if (!uniforms || !states || !phi) return fail(-1, "invalid arguments");
    if (!(phi[0] > 0 && phi[1] > 0)) return fail(-1, "invalid spot parameters");
    std::vector<double> D;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        std::vector<double> L, G;
        int rc = eval_explicit(xyz, P, T, data, bg, 1, phi, 1, model, MODE_DELTA, L, G, &D);
        if (rc) return rc;
        if (logL) *logL = L[0];
    }
    if (delta) memcpy(delta, D.data(), 3 * (size_t)T * sizeof(double));
    double *dd = nullptr, *du = nullptr;
    uint8_t *ds = nullptr;
    hipError_t e = hipMalloc(&dd, D.size() * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&du, (size_t)T * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ds, (size_t)T);
    if (e == hipSuccess) e = hipMemcpy(dd, D.data(), D.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(du, uniforms, (size_t)T * sizeof(double), hipMemcpyHostToDevice);
    int rc = 0;
    if (e == hipSuccess) {
        rc = rfba3b_backward_sample_device(dd, 1, T, model, du, ds, nullptr);
        if (!rc) e = hipMemcpy(states, ds, (size_t)T, hipMemcpyDeviceToHost);
    }
    hipFree(dd);
    hipFree(du);
    hipFree(ds);
    if (rc) return rc;
    if (e != hipSuccess) return fail(-3, "backward sampling: %s", hipGetErrorString(e));
    return 0;
}

// The reference's MT19937 stream (mt19937.h:48-63 over mersenne.cpp): seeded
// by the standard MT19937 initialisation, 37 outputs discarded (RandomInit),
// then each draw is the 32-bit output times 2^-32.  std::mt19937 implements
// the same generator and initialisation.
extern "C" int rfba3b_mt19937_uniforms(int seed, unsigned long long skip, size_t n, double *out) {
    if (!out && n) return fail(-1, "NULL output");
    std::mt19937 g((uint32_t)seed);
    g.discard(37 + skip);
    for (size_t i = 0; i < n; i++) out[i] = (double)g() * (1.0 / (65536.0 * 65536.0));
    return 0;
}

// ---------------------------------------------------------------------------
// FFBS Hessian of the resident functor: sampled_background_spot_hessian_ffbs
// (multispot5.cc:902-982) for the handle's spot data.

extern "C" int rfba3b_functor_hessian_ffbs(rfba3b_functor *f, const double *phi, int bs, const double *uniforms,
                                           double *hess, double *deriv, uint8_t *paths, void *stream) {
    if (!f) return fail(-1, "handle is NULL");
    if (!phi || !uniforms || !hess || bs <= 0) return fail(-1, "invalid arguments");
    if (!(phi[0] > 0 && phi[1] > 0)) return fail(-1, "invalid spot parameters");
    std::lock_guard<std::mutex> hl(f->mu);
    const int S = f->S, T = f->T, X = f->X, Y = f->Y, Z = f->Z;
    const size_t Pd = (size_t)X * Y * Z;
    const size_t nchain = (size_t)S * bs;
    if (nchain > (size_t)1 << 24 || nchain * T > (size_t)1 << 31) return fail(-1, "S x bs_iterations x T too large");
    const int nbas = (int)((Pd + 255) / 256);
    const int ngrp = (T + kHessFrames - 1) / kHessFrames;
    hipStream_t st = (hipStream_t)stream;
    std::vector<double> parts((size_t)S * T * kHessParts), fixed((size_t)nbas * kHessUpper);
    std::vector<uint8_t> path(nchain * T);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        hipError_t e = hipSuccess;
        auto al = [&](void **p, size_t bytes) {
            if (e == hipSuccess && !*p) e = hipMalloc(p, std::max<size_t>(bytes, 16));
        };
        al((void **)&f->d_basis, Pd * kHessBasis * sizeof(float));
        al((void **)&f->d_fixed, (size_t)nbas * kHessUpper * sizeof(double));
        al((void **)&f->d_parts, (size_t)S * T * kHessParts * sizeof(double));
        al((void **)&f->d_delta, (size_t)S * T * 3 * sizeof(double));
        if (e == hipSuccess && nchain > f->paths_cap) {
            hipFree(f->d_unif);
            hipFree(f->d_paths);
            f->d_unif = nullptr;
            f->d_paths = nullptr;
            f->paths_cap = 0;
            al((void **)&f->d_unif, nchain * T * sizeof(double));
            al((void **)&f->d_paths, nchain * T);
            if (e == hipSuccess) f->paths_cap = nchain;
        }
        if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? -2 : -3, "Hessian buffers: %s", hipGetErrorString(e));
        // forward messages of the spot over each background (the handle's
        // planes are data - bg), then bs backward samplings per background
        for (int c = 0; c < 5; c++) ((double *)f->h_stage)[c] = phi[c];
        int rc = functor_sums(f, 1, st);
        if (!rc) {
            HmmParams H = functor_hmm(f, S);
            H.delta = f->d_delta;
            rc = launch_hmm(MODE_DELTA, H, T, st);
        }
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(f->d_unif, uniforms, nchain * T * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(b3_backward_sample, dim3((unsigned)((nchain + 255) / 256)), dim3(256), 0, st, f->d_delta,
                           (int)nchain, bs, T, log_a(&f->model), f->d_unif, f->d_paths);
        HIP_TRY(hipGetLastError());
        // spot table and the per-frame parts
        HessParams Hp;
        memset(&Hp, 0, sizeof Hp);
        Hp.block = f->d_block; Hp.model = f->d_model; Hp.basis = f->d_basis; Hp.fixed = f->d_fixed; Hp.R = f->d_R;
        Hp.S = S; Hp.T = T; Hp.parts = f->d_parts;
        hipLaunchKernelGGL(b3_hess_basis, dim3(nbas), dim3(256), 0, st, Hp);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(b3_hess_parts, dim3(S * ngrp), dim3(256), 0, st, Hp);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(parts.data(), f->d_parts, parts.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(fixed.data(), f->d_fixed, fixed.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(path.data(), f->d_paths, path.size(), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    if (paths) memcpy(paths, path.data(), path.size());
    // Frame-independent sums of d_i d_j, in workgroup order; the index of
    // entry (i, j), i <= j, in the upper-triangle order.
    double F[kHessUpper] = {0};
    for (int w = 0; w < nbas; w++)
        for (int k = 0; k < kHessUpper; k++) F[k] += fixed[(size_t)w * kHessUpper + k];
    int U[5][5];
    for (int i = 0, k = 0; i < 5; i++)
        for (int j = i; j < 5; j++) U[i][j] = U[j][i] = k++;
    // hess_t = sum e H - d d^T (the reference's entry [4][3] of H is never
    // set, storm.h:143: taken as 0 in reference mode), deriv_t = sum e d;
    // per path H = sum over on frames, D likewise; sum (H + D D^T) and sum D
    // over backgrounds and paths (hess_handler, multispot5.cc:839-887).
    std::vector<double> Ht(25 * (size_t)T), Dt(5 * (size_t)T);
    double sh[25] = {0}, sd[5] = {0};
    for (int s = 0; s < S; s++) {
        for (int t = 0; t < T; t++) {
            const double *pt = &parts[((size_t)s * T + t) * kHessParts];
            for (int i = 0; i < 5; i++)
                for (int j = 0; j < 5; j++) {
                    const double eH = (i == 4 && j == 3 && !f->model.z_exact) ? 0.0 : pt[U[i][j]];
                    Ht[25 * (size_t)t + 5 * i + j] = eH - F[U[i][j]];
                }
            for (int i = 0; i < 5; i++) Dt[5 * (size_t)t + i] = pt[kHessUpper + i];
        }
        for (int r = 0; r < bs; r++) {
            const uint8_t *pa = &path[((size_t)s * bs + r) * T];
            double H[25] = {0}, D[5] = {0};
            for (int t = 0; t < T; t++)
                if (pa[t] == 0) {
                    for (int i = 0; i < 25; i++) H[i] += Ht[25 * (size_t)t + i];
                    for (int i = 0; i < 5; i++) D[i] += Dt[5 * (size_t)t + i];
                }
            for (int i = 0; i < 5; i++)
                for (int j = 0; j < 5; j++) sh[5 * i + j] += H[5 * i + j] + D[i] * D[j];
            for (int i = 0; i < 5; i++) sd[i] += D[i];
        }
    }
    // :960-975: normalise, subtract the mean outer product, add the priors
    const double norm = (double)bs * S * f->model.variance;
    for (int i = 0; i < 25; i++) sh[i] /= norm;
    for (int i = 0; i < 5; i++) sd[i] /= norm;
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) sh[5 * i + j] -= sd[i] * sd[j];
    const rfba3b_model &m = f->model;
    auto hlln = [](double x, double mu, double sg) { return (1 + (ln0(x) - mu - 1) / (sg * sg)) / (x * x); };
    sh[0] += hlln(phi[0], m.mu_brightness, m.sigma_brightness);
    sh[6] += hlln(phi[1], m.mu_blur, m.sigma_blur);
    memcpy(hess, sh, sizeof sh);
    if (deriv) {
        sd[0] += dlln(phi[0], m.mu_brightness, m.sigma_brightness);
        sd[1] += dlln(phi[1], m.mu_blur, m.sigma_blur);
        memcpy(deriv, sd, sizeof sd);
    }
    return 0;
}
