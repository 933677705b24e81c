/*
 * rfba_srrf.h -- C ABI of the MI355X (gfx950) 3D SRRF radiality library
 * (librfba_srrf.so).
 *
 * Three entry points keep the reference's DLL contract exactly (same names,
 * argument meaning, struct layout and INI file semantics):
 *
 *   SRRFGradientCalc     replaces RFBA Main/GradientCalc/GradientCalc.h:4
 *                        (GradientCalc.cpp:169-173, tiling :120-163)
 *   SRRFGetMag           replaces GradientCalc.h:5 (GradientCalc.cpp:176-178)
 *   SRRFGradientCalcGPU  replaces RFBA Main/GradientCalc/kernel.cuh:26
 *                        (kernel.cu:560-657)
 *
 * The rfba_srrf_* entry points are additions: explicit sampling sets,
 * uint16 input, device-resident buffers, and the batched form.
 *
 * Conventions: every volume is x-fastest, [t][z][y][x] for series and
 * [z][y][x] for SR outputs, exactly as getIdx3 (kernel.cu:36-40).
 * Functions return RFBA_OK (0) on success, a negative rfba_status otherwise;
 * rfba_last_error() describes the last failure of the calling thread.
 */
#ifndef RFBA_SRRF_H
#define RFBA_SRRF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The reference never reports errors (it always returns 0);
 * this library returns 0 on success exactly as the reference does and a
 * negative code on failure. */
typedef enum rfba_status {
    RFBA_OK = 0,
    RFBA_EINVAL = -1,  /* invalid argument / configuration */
    RFBA_ENOMEM = -2,  /* host or device allocation failed */
    RFBA_EDEVICE = -3, /* HIP runtime error (launch, copy, sync) */
    RFBA_ENODEV = -4   /* no usable gfx950 device */
} rfba_status;

/* Byte-for-byte the reference's SRRFConf (kernel.cuh:3-24): 17 x 4 bytes. */
typedef struct SRRFConf {
    float RingRadius;   /* ring radius in SR pixels (kernel.cu:220-222) */
    int Magnification;  /* M */
    int AxesinRing;     /* nRing = 2*AxesinRing, nRing^2 samples */
    int MainRange;      /* gradient box half-width along the derivative axis */
    int XRange;
    int YRange;
    int ZRange;
    int Method;         /* 0 TRM, 1 TRA, 2 TRPPM, 3 TRAC */
    int ITC;            /* TRAC: integrate lag times (binning levels) */
    int TRACOrder;      /* 2, 3 or 4 */
    int RemoveCon;      /* 1 applies max(radiality, 0) (naming as reference) */
    int Renormalize;    /* 1: 0.5 + radiality/2 */
    int GradientSmooth; /* 0: central difference, 1: box sums */
    int IntensityW;     /* intensity weighting */
    int blockSizeX;     /* tiling of SRRFGradientCalc (default 16) */
    int blockSizeY;
    int blockSizeZ;
} SRRFConf;

/* ---- reference-compatible entry points ------------------------------- */

/* Whole-volume SRRF with the reference's implicit configuration: reads
 * ".\SRRF.ini" (here "./SRRF.ini") on every call, defaults when absent
 * (GradientCalc.cpp:15-42), splits the volume into blockSize^3 tiles with a
 * 2-voxel border and stitches them with the reference's put-back rule
 * (GradientCalc.cpp:47-116): the global low band [0, 2M) of every axis is
 * left untouched, later tiles overwrite earlier ones.
 * inImage: float[T][Z][Y][X] host; imageDim = {X, Y, Z, T};
 * outImage: caller-allocated float[Z*M][Y*M][X*M] host. */
int SRRFGradientCalc(float *inImage, int *imageDim, float *outImage);

/* Magnification from "./SRRF.ini" ([SRRF] Magnification, default 5). */
int SRRFGetMag(void);

/* One stack, no tiling, explicit configuration, reference ring sampling. */
int SRRFGradientCalcGPU(float *inImage, int *imageDim, float *outImage,
                        SRRFConf *conf);

/* ---- configuration helpers ------------------------------------------- */

/* Parse an SRRF.ini-style file with the reference's section/key names and
 * defaults (GradientCalc.cpp:15-42). path == NULL reads "./SRRF.ini".
 * A missing file yields the defaults and RFBA_OK. */
int rfba_srrf_load_ini(const char *path, SRRFConf *conf);

/* Fill conf with the reference defaults (GradientCalc.cpp:19-41). */
void rfba_srrf_default_conf(SRRFConf *conf);

/* ---- sampling sets ------------------------------------------------------
 * A sampling set is an ordered list of n offsets (x, y, z interleaved, in SR
 * pixels) added to the SR voxel centre, plus the divisor applied to the
 * signed sum of the per-sample terms.  Reference ring mode: n = nRing^2,
 * divisor = nRing (kernel.cu:23-33, 244). */
typedef struct rfba_srrf_samples {
    int n;              /* number of samples, 1..RFBA_SRRF_MAX_SAMPLES */
    float divisor;      /* radiality = signed sum / divisor */
    const float *xyz;   /* host pointer, 3*n floats */
} rfba_srrf_samples;

#define RFBA_SRRF_MAX_SAMPLES 1024

/* Reference ring: nRing = 2*axes_in_ring; writes 3*nRing^2 floats in the
 * reference's order (idx = i1*nRing + i2) and returns nRing^2 (or < 0). */
int rfba_srrf_build_ring(int axes_in_ring, float radius, float *xyz_out);

/* Axis-aligned (kind 6) or 3x3x3-neighbourhood (kind 26) direction sets
 * scaled to `radius`; offsets are computed in double and rounded once.
 * Listed by |x| (largest first), then x, y, z: samples sharing an x offset,
 * and those sharing an (x, y) offset, are consecutive, and x = -radius and
 * x = +radius are neighbours.  Returns n. */
int rfba_srrf_build_dirs(int kind, float radius, float *xyz_out);

/* ---- explicit single-stack entry points --------------------------------
 * dtype: RFBA_U16 (uint16 frames) or RFBA_F32 (float frames).
 * samples == NULL selects the reference ring from conf. */
#define RFBA_U16 1
#define RFBA_F32 2

/* Host buffers in and out (includes H2D/D2H). */
int rfba_srrf_stack(const void *in, int dtype, const int dims[4],
                    const SRRFConf *conf, const rfba_srrf_samples *samples,
                    float *out);

/* Device-resident buffers: d_in and d_out are device pointers; the call is
 * asynchronous on `stream` (hipStream_t, NULL = default stream) unless
 * kernel_ms != NULL, in which case it synchronises and reports the time
 * between HIP events recorded around the kernel launches on that stream. */
int rfba_srrf_stack_device(const void *d_in, int dtype, const int dims[4],
                           const SRRFConf *conf,
                           const rfba_srrf_samples *samples, float *d_out,
                           void *stream, float *kernel_ms);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc).  Same
 * inputs, outputs and results as rfba_srrf_stack_device, but the run's
 * sample and tile tables, queue counter and TRAC / TRPPM frame series live
 * in a static 256 MiB device array of the library (one per GPU, held until
 * the process ends): no hipMalloc or hipFree, also on the first call.
 * Limits: at most 1,048,576 SR tiles (16 x 16 x 4..16 voxels each), and a
 * TRAC / TRPPM series area of 239.9 MiB, of which each concurrently running
 * workgroup takes frames x 4..16 x 256 floats: when the occupancy grid needs
 * more, fewer workgroups run (same results, longer time); sizes beyond
 * either limit return RFBA_EINVAL before any device work.  Synchronous on
 * `stream`; calls are serialised within the process (a second process has
 * its own array), and the array keeps the last call's contents. */
int rfba_srrf_stack_device_noalloc(const void *d_in, int dtype,
                                   const int dims[4], const SRRFConf *conf,
                                   const rfba_srrf_samples *samples,
                                   float *d_out, void *stream);

/* ---- batched form over uneven blocks -----------------------------------
 * A block is a sub-volume (origin, extent) of a source series with its own
 * magnification, or a whole stack; every block is computed exactly as if it
 * were passed alone to rfba_srrf_stack (its own clamp-to-edge borders), and
 * its SR result goes to its own output buffer.  One persistent work-queue
 * launch covers all blocks. */
typedef struct rfba_srrf_block {
    const void *d_src;  /* device pointer to the source series */
    int src_dims[4];    /* X, Y, Z, T of the source series */
    int origin[3];      /* block origin inside the source (low-res voxels) */
    int dims[3];        /* block extent (low-res voxels) */
    int frames;         /* number of frames used (<= src_dims[3]) */
    int mag;            /* magnification of this block */
    float *d_out;       /* device float[dims.z*mag][dims.y*mag][dims.x*mag] */
} rfba_srrf_block;

typedef struct rfba_srrf_batch rfba_srrf_batch; /* opaque plan */

/* Build a plan for n blocks that share dtype, conf (Method, smoothing, ...)
 * and samples.  Blocks are grouped by magnification; each group gets its own
 * tile depth, LDS layout and persistent launch (the launches run back to back
 * on one stream, one hipGraph).  The plan owns its device tables and is
 * reusable while the block list is unchanged. */
int rfba_srrf_batch_create(const rfba_srrf_block *blocks, int n, int dtype,
                           const SRRFConf *conf,
                           const rfba_srrf_samples *samples,
                           rfba_srrf_batch **plan);
/* Run: one persistent launch per magnification group over all its blocks
 * (captured into a hipGraph on the first call when use_graph != 0, replayed
 * afterwards).  With kernel_ms != NULL the call waits for the launches,
 * stores their time and returns the kernel's status; with kernel_ms == NULL
 * it returns after enqueueing, and rfba_srrf_batch_status reports how the
 * run ended. */
int rfba_srrf_batch_run(rfba_srrf_batch *plan, void *stream, int use_graph,
                        float *kernel_ms);
/* Reference form for comparison: one launch per block on the same stream
 * (kernel_ms as for rfba_srrf_batch_run). */
int rfba_srrf_batch_run_per_block(rfba_srrf_batch *plan, void *stream,
                                  float *kernel_ms);
/* Waits for the plan's last run and returns RFBA_OK, or RFBA_EDEVICE when a
 * kernel of that run found a tile larger than the planned LDS extents (its
 * outputs are then incomplete). */
int rfba_srrf_batch_status(rfba_srrf_batch *plan);
/* Autotuner (no reference counterpart): per magnification group, times
 * every feasible tile depth (4, 8, 16 SR voxels along z) on a sub-plan of
 * every sample_stride-th block of the group (0: about 32 blocks), keeps the
 * fastest and rebuilds the group for it.  *tz and ms_by_tz[3] (sample times
 * at depth 4, 8, 16; 0 = not feasible) describe the group with the most
 * work; either may be NULL.  Writes the sampled blocks' outputs.
 * Synchronous on stream. */
int rfba_srrf_batch_autotune(rfba_srrf_batch *plan, void *stream,
                             int sample_stride, int *tz, float *ms_by_tz);
/* Tile depth of the group with the most work. */
int rfba_srrf_batch_tile_depth(const rfba_srrf_batch *plan);
/* Magnification and tile depth of each group (largest work first), up to
 * max_groups entries; returns the number of groups. */
int rfba_srrf_batch_groups(const rfba_srrf_batch *plan, int *mags,
                           int *tile_depths, int max_groups);
void rfba_srrf_batch_destroy(rfba_srrf_batch *plan);

/* ---- runtime helpers ---------------------------------------------------- */
const char *rfba_last_error(void);
int rfba_device_count(void);
int rfba_malloc_device(void **ptr, size_t bytes);
int rfba_free_device(void *ptr);
int rfba_memcpy_h2d(void *dst, const void *src, size_t bytes);
int rfba_memcpy_d2h(void *dst, const void *src, size_t bytes);
int rfba_memset_device(void *dst, int value, size_t bytes);
int rfba_device_synchronize(void);
/* Plain device-to-device copy kernel of `bytes` (multiple of 16): the
 * reference point used to calibrate HBM byte counters. */
int rfba_copy_probe(void *d_dst, const void *d_src, size_t bytes, float *ms);

/* ---- caller post-processing and the SRRF.raw file ------------------------
 * The reference's caller clamps the SR volume (NaN -> 0, negative -> 0) and
 * keeps its maximum, starting from 1 (multispot5_place_methods.cc:260-269);
 * then writes it as raw float32, x fastest, to SRRF.raw (:270-274) and reads
 * it back in split mode (:284-289). */
int rfba_srrf_postprocess_device(float *d_out, size_t n, void *stream,
                                 float *fmax);
int rfba_srrf_write_raw(const char *path, const float *data, size_t n);
/* Fails (RFBA_EINVAL) when the file holds fewer than n floats; the
 * reference's fread would leave the tail of the buffer unread. */
int rfba_srrf_read_raw(const char *path, float *data, size_t n);

/* SRRF-guided spot placement (multispot5_place_methods.cc:293-341): every
 * low-res pixel (x, y, z) of the region (pixels [npix][3]) expands to its
 * mag^3 SR pixels (x offset outermost, z innermost); their values are read
 * from the device SR volume d_sr [Z*mag][Y*mag][X*mag] (dims = X, Y, Z);
 * spots are then placed by rejection sampling against the region's maximum
 * with the reference's MT19937 stream for `seed` (get_defaults, :69-71):
 * spots[k] = (intensity_mode, blur_mode, x, y, z), positions in low-res
 * pixels.  Returns the number of spots placed (0 when the maximum is
 * negative, as the reference returns none) or a negative status;
 * *draws_used (may be NULL) = generator draws consumed. */
int rfba_srrf_place_spots(const float *d_sr, const int *dims, int mag,
                          const int *pixels, int npix, int num_spots,
                          double intensity_mode, double blur_mode, int seed,
                          double *spots, unsigned long long *draws_used);
/* Release the library's cached device scratch (kept between calls). */
void rfba_release_cache(void);

#ifdef __cplusplus
}
#endif
#endif /* RFBA_SRRF_H */
