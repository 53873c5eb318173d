/*
 * ddm_oracle.c -- CPU restatement of the multi-scale DDM analysis path of
 * zhangyuyi99/cuda_multiDDM (tile -> 2-D R2C FFT -> lag difference ->
 * masked azimuthal average).
 *
 * THIS IS TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (FP64
 * build, functions suffixed _f64) or as the timed CPU baseline (FP32
 * multithreaded build, suffixed _f32).  The product path never calls it.
 *
 * What each function restates (reference file:line):
 *   ddmo_pixel / ddmo_tile  parseBufferScalePow2, src/DDM_kernel.cuh:43-78
 *                           (ROI pixel (x_off + tx*s + lx, y_off + ty*s + ly),
 *                           tile j = ty*(M/s) + tx, channel 0 of a bpp-byte
 *                           pixel, src/DDM.cu:188).  16-bit formats are read
 *                           whole (the reference truncates to the low byte).
 *   rfft2                   cufftExecR2C, src/DDM.cu:189-192 with the plan of
 *                           src/DDM.cu:654-677: unnormalised forward 2-D R2C,
 *                           out[ky][kx], kx <= s/2.
 *   lag_block               processFFT, src/DDM_kernel.cuh:85-101 and the
 *                           slot choice of analyseChunk, src/DDM.cu:222-262,
 *                           restated as pairs (t, t+tau) with both frames
 *                           inside the video.
 *   ddmo_mask               buildAzimuthMask, src/azimuthal_average.cu:229-283
 *                           (FP32 arithmetic, operation for operation), with
 *                           the q radius of src/DDM.cu:101-105.
 *   average (in ddmo_run)   analyseFFTDevice, src/azimuthal_average.cu:292-346:
 *                           I = count > 0 ? 2 * sum_mask(acc) / count * norm : 0
 *                           with norm = 1/N (src/DDM.cu:92, 875).
 *   ddmo_sorted_params      the in-place sorts of src/DDM.cu:319-321.
 *   ddmo_check              the checks of src/DDM.cu:327-350, 459-463, with
 *                           the two hard limits (tau_max <= chunk,
 *                           lambda_max <= s_min) lifted.
 *
 * Pinning (tests/test_oracle.py): the FFT against a brute-force DFT and
 * numpy.fft.rfft2 fixtures, closed forms (constant tile, delta, plane
 * waves, Parseval); the lag sums against closed forms (identical frames,
 * alternating frames, linear ramps) and brute force; the mask against exact
 * integer-radius arithmetic away from FP32 rounding boundaries; the average
 * against constant accumulators.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define DDMO_EINVAL (-1)
#define DDMO_ENOMEM (-2)
#define DDMO_EROI (-3)
#define DDMO_ESCALE (-4)
#define DDMO_ELAG (-5)
#define DDMO_ELAMBDA (-6)

enum { DDMO_PIX_U8 = 0, DDMO_PIX_U16LE = 1, DDMO_PIX_U16BE = 2 };
enum { DDMO_NORM_FRAMES = 0, DDMO_NORM_PAIRS = 1 };

typedef struct {
    int width, height;   /* frame size in pixels */
    int bpp;             /* bytes per pixel (channel stride) */
    int pixfmt;          /* DDMO_PIX_* */
    int x_off, y_off;    /* ROI origin */
    int n_frames;
    const int *scales;
    int scale_count;
    const int *lags;
    int lag_count;
    const float *lambdas;
    int lambda_count;
    float tolerance;     /* ring [q, tol*q] */
    int norm_mode;       /* DDMO_NORM_* */
    int nthreads;
    int only_scale;      /* >= 0: process only this index of the descending scale list */
    int max_tiles;       /* > 0: process only the first max_tiles tiles of each scale */
    const int *tile_list;  /* tile_list_count > 0: process only these tile indices of each scale */
    int tile_list_count;   /* (indices past a scale's tile count are ignored) */
} ddmo_params;

static int ilog2_exact(int n) {
    if (n <= 0 || (n & (n - 1))) return -1;
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

static int cmp_int_asc(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}
static int cmp_int_desc(const void *a, const void *b) { return cmp_int_asc(b, a); }
static int cmp_float_asc(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

/* Sorted copies: lags ascending, lambdas ascending, scales descending. */
static int ddmo_sorted_params(const ddmo_params *p, int **scales, int **lags, float **lambdas) {
    *scales = (int *)malloc(sizeof(int) * (size_t)p->scale_count);
    *lags = (int *)malloc(sizeof(int) * (size_t)p->lag_count);
    *lambdas = (float *)malloc(sizeof(float) * (size_t)p->lambda_count);
    if (!*scales || !*lags || !*lambdas) {
        free(*scales); free(*lags); free(*lambdas);
        return DDMO_ENOMEM;
    }
    memcpy(*scales, p->scales, sizeof(int) * (size_t)p->scale_count);
    memcpy(*lags, p->lags, sizeof(int) * (size_t)p->lag_count);
    memcpy(*lambdas, p->lambdas, sizeof(float) * (size_t)p->lambda_count);
    qsort(*scales, (size_t)p->scale_count, sizeof(int), cmp_int_desc);
    qsort(*lags, (size_t)p->lag_count, sizeof(int), cmp_int_asc);
    qsort(*lambdas, (size_t)p->lambda_count, sizeof(float), cmp_float_asc);
    return 0;
}

static int ddmo_check(const ddmo_params *p) {
    if (p->scale_count < 1 || p->lag_count < 1 || p->lambda_count < 1 || p->n_frames < 1)
        return DDMO_EINVAL;
    if (p->bpp < 1) return DDMO_EINVAL;
    if ((p->pixfmt == DDMO_PIX_U16LE || p->pixfmt == DDMO_PIX_U16BE) && p->bpp < 2) return DDMO_EINVAL;
    int M = 0;
    for (int i = 0; i < p->scale_count; ++i) {
        if (ilog2_exact(p->scales[i]) < 0) return DDMO_ESCALE;
        for (int j = 0; j < i; ++j)
            if (p->scales[j] == p->scales[i]) return DDMO_ESCALE;
        if (p->scales[i] > M) M = p->scales[i];
    }
    for (int i = 0; i < p->lag_count; ++i) {
        if (p->lags[i] < 0) return DDMO_ELAG;
        for (int j = 0; j < i; ++j)
            if (p->lags[j] == p->lags[i]) return DDMO_ELAG;
    }
    for (int i = 0; i < p->lambda_count; ++i) {
        if (!(p->lambdas[i] >= 0.0f)) return DDMO_ELAMBDA;
        for (int j = 0; j < i; ++j)
            if (p->lambdas[j] == p->lambdas[i]) return DDMO_ELAMBDA;
    }
    if (p->x_off < 0 || p->y_off < 0 || M + p->x_off > p->width || M + p->y_off > p->height)
        return DDMO_EROI;
    return 0;
}

static double ddmo_norm(const ddmo_params *p, int tau) {
    if (p->norm_mode == DDMO_NORM_PAIRS) {
        long pairs = (long)p->n_frames - tau;
        return pairs > 0 ? 1.0 / (double)pairs : 0.0;
    }
    return 1.0 / (double)p->n_frames;
}

/* Pixel value at (x, y) of frame f: channel 0 of a bpp-byte pixel. */
static double ddmo_pixel(const ddmo_params *p, const uint8_t *frames, int f, int x, int y) {
    size_t off = (((size_t)f * p->height + (size_t)y) * p->width + (size_t)x) * (size_t)p->bpp;
    switch (p->pixfmt) {
    case DDMO_PIX_U16LE: return (double)((unsigned)frames[off] | ((unsigned)frames[off + 1] << 8));
    case DDMO_PIX_U16BE: return (double)(((unsigned)frames[off] << 8) | (unsigned)frames[off + 1]);
    default: return (double)frames[off];
    }
}

/* Tile `tile` of scale s (largest scale M) of frame f into out[ly][lx]. */
void ddmo_tile(const ddmo_params *p, const uint8_t *frames, int M, int s, int tile, int f, double *out) {
    const int tps = M / s;
    const int ty = tile / tps, tx = tile % tps;
    for (int ly = 0; ly < s; ++ly)
        for (int lx = 0; lx < s; ++lx)
            out[(size_t)ly * s + lx] =
                ddmo_pixel(p, frames, f, p->x_off + tx * s + lx, p->y_off + ty * s + ly);
}

/*
 * q-ring mask of scale s for one lambda, evaluated in FP32 exactly as
 * buildAzimuthMask does (src/azimuthal_average.cu:229-283), with
 * q = float(s) / lambda (src/DDM.cu:102).  mask[ky][kx] (kx <= s/2).
 * Returns the member count.  This translation unit is compiled with
 * -ffp-contract=off so no step is fused.
 */
int ddmo_mask(int s, float lambda, float tol, uint8_t *mask) {
    volatile float q = (float)s / lambda;
    volatile float q2 = q * q;
    volatile float tol2 = tol * tol;
    const int w = s, h = s, half_w = w / 2, half_h = h / 2;
    int count = 0;
    for (int x = 0; x < w / 2 + 1; ++x) {
        for (int y = 0; y < h; ++y) {
            int xs = (x + half_w) % w - half_w;
            int ys = (y + half_h) % h - half_h;
            volatile float r2 = (float)(xs * xs + ys * ys);
            volatile float ratio = r2 / q2;
            int px = (1.0f <= ratio) && (ratio <= tol2);
            count += px;
            mask[(size_t)y * (w / 2 + 1) + x] = (uint8_t)px;
        }
    }
    return count;
}

/* Brute-force O(s^4) 2-D DFT (FP64, exact twiddle index reduction). */
int ddmo_dft2_brute(const double *in, int s, double *out) {
    if (s < 1) return DDMO_EINVAL;
    double *c = (double *)malloc(sizeof(double) * (size_t)s);
    double *sn = (double *)malloc(sizeof(double) * (size_t)s);
    if (!c || !sn) { free(c); free(sn); return DDMO_ENOMEM; }
    for (int k = 0; k < s; ++k) {
        c[k] = cos(-2.0 * M_PI * k / s);
        sn[k] = sin(-2.0 * M_PI * k / s);
    }
    const int h = s / 2 + 1;
    for (int ky = 0; ky < s; ++ky)
        for (int kx = 0; kx < h; ++kx) {
            double re = 0, im = 0;
            for (int y = 0; y < s; ++y)
                for (int x = 0; x < s; ++x) {
                    int m = (int)(((long)ky * y + (long)kx * x) % s);
                    double v = in[(size_t)y * s + x];
                    re += v * c[m];
                    im += v * sn[m];
                }
            out[2 * ((size_t)ky * h + kx)] = re;
            out[2 * ((size_t)ky * h + kx) + 1] = im;
        }
    free(c);
    free(sn);
    return 0;
}

/* FP64 oracle */
#define REAL double
#define SUF(x) x##_f64
#include "ddm_oracle_impl.h"
#undef REAL
#undef SUF

/* FP32 multithreaded CPU baseline */
#define REAL float
#define SUF(x) x##_f32
#include "ddm_oracle_impl.h"
#undef REAL
#undef SUF

/* Mask count helper for tests: fills counts[q] for the sorted lambdas. */
int ddmo_mask_counts(int s, const float *lambdas, int q_count, float tol, int *counts) {
    uint8_t *m = (uint8_t *)malloc((size_t)s * (s / 2 + 1));
    if (!m) return DDMO_ENOMEM;
    for (int q = 0; q < q_count; ++q) counts[q] = ddmo_mask(s, lambdas[q], tol, m);
    free(m);
    return 0;
}
