/*
 * rfba_prep.h -- C ABI of the MI355X (gfx950) input library (librfba_prep.so):
 * RFBA's image loading and normalisation, the step ahead of both the SRRF
 * stage and the 3B fits.
 *
 * Reference:
 *   img_load3D                   RFBA Main/storm_imagery.cc:185-254
 *     (one multi-page 8- or 16-bit TIFF = one 3D stack, pages = z)
 *   load_and_preprocess_images   storm_imagery.cc:269-287
 *   preprocess_image             storm_imagery_no_load.cc:44-79
 *     (per z-slice high-pass: the slice minus its Gaussian blur,
 *      libcvd convolveGaussian_fir with sigma = preprocess.lpf and 3 sigmas,
 *      on the interior; the one-pixel border is 0; skipped when
 *      preprocess.skip)
 *   mean_and_variance            storm_imagery_no_load.cc:87-110
 *   load_and_normalize_images    storm_imagery.cc:358-416
 *     (divide by the standard deviation, or (x - subtract) / divide with
 *      preprocess.fixed_scaling)
 * Config keys: multispot5.cfg:49-52.  Return 0 on success, negative on error
 * (rfba_prep_last_error): -1 invalid argument, -2 out of memory, -3 device
 * or file error, -4 no gfx950 device.
 */
#ifndef RFBA_PREP_H
#define RFBA_PREP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- TIFF (host) --------------------------------------------------------
 * Uncompressed, one sample per pixel, 8 or 16 bits, either byte order, any
 * strip layout; every page the same size. */
typedef struct rfba_tiff_info {
    int width, height, pages, bits;
} rfba_tiff_info;

/* Size and sample depth of a file without reading its pixels. */
int rfba_tiff_info_read(const char *path, rfba_tiff_info *info);
/* The pixels as [pages][height][width] of uint8 or uint16 (native order)
 * into out (capacity cap bytes); info may be NULL. */
int rfba_tiff_load(const char *path, void *out, size_t cap, rfba_tiff_info *info);
/* Writes an uncompressed little-endian (big_endian = 0) or big-endian
 * multi-page TIFF with rows_per_strip rows per strip (0 = whole page). */
int rfba_tiff_write(const char *path, const void *px, int width, int height, int pages, int bits, int big_endian,
                    int rows_per_strip);

/* ---- normalisation (device) --------------------------------------------- */
typedef struct rfba_prep_conf {
    int skip;              /* preprocess.skip: no high-pass */
    float lpf;             /* preprocess.lpf: sigma of the blur (5) */
    int fixed_scaling;     /* preprocess.fixed_scaling (requires skip) */
    double subtract;       /* preprocess.fixed_scaling.subtract */
    double divide;         /* preprocess.fixed_scaling.divide */
} rfba_prep_conf;

enum { RFBA_PREP_U8 = 8, RFBA_PREP_U16 = 16, RFBA_PREP_F32 = 32 };

/* convolveGaussian's FIR (libcvd convolution.h:400-416): half-width
 * ksize = ceil(sigmas * sigma), kernel[i-1] = exp(-i^2 / (2 sigma^2)) as
 * float divided by 2 sum + 1, centre weight factor = 1 / (2 sum + 1). */
int rfba_prep_gaussian_kernel(double sigma, double sigmas, float *kernel, int cap, int *ksize, double *factor);

/* load_and_normalize_images on a device-resident series: d_raw [T][Z][Y][X]
 * of type `type`, dims = {X, Y, Z, T}.  Writes d_out [T][Z][Y][X] float and
 * stats = {mean, variance} before scaling and {mean, variance} after
 * (mean_and_variance, as floats).  Synchronous; kernel_ms may be NULL. */
int rfba_prep_normalize_device(const void *d_raw, int type, const int *dims, const rfba_prep_conf *conf,
                               float *d_out, double *stats, void *stream, float *kernel_ms);

/* The whole step from files: n TIFF stacks (one per frame, all the same
 * size) read on the host, copied to the device through pinned buffers and
 * normalised into d_out [n][Z][Y][X] float (capacity cap floats); dims gets
 * {X, Y, Z, n}. */
int rfba_prep_load_and_normalize(const char *const *paths, int n, const rfba_prep_conf *conf, float *d_out,
                                 size_t cap, int *dims, double *stats, void *stream);

/* Device memory for callers without a HIP runtime of their own. */
int rfba_prep_malloc(void **p, size_t bytes);
int rfba_prep_free(void *p);
int rfba_prep_memcpy(void *dst, const void *src, size_t bytes, int kind); /* 1 H2D, 2 D2H, 3 D2D */

const char *rfba_prep_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* RFBA_PREP_H */
