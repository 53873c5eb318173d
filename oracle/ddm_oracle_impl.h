/*
 * Precision-generic body of the CPU restatement of the multi-scale DDM
 * analysis path.  Included twice by ddm_oracle.c: once with REAL=double
 * (the FP64 oracle) and once with REAL=float (the multithreaded FP32 CPU
 * baseline).  TEST INFRASTRUCTURE ONLY -- see the header of ddm_oracle.c.
 *
 * Requires the includer to define REAL and SUF(name).
 */

/* ---------------------------------------------------------------------------
 * Complex radix-2 FFT (iterative, decimation in time, forward sign -1).
 * Twiddles are computed in double with cos/sin of 2*pi*k/n and rounded to
 * REAL once, so the FP32 baseline uses correctly rounded twiddles.
 * ------------------------------------------------------------------------- */
typedef struct { REAL re, im; } SUF(cplx);

typedef struct {
    int n;
    int log2n;
    SUF(cplx) *tw;   /* tw[k] = exp(-2 pi i k / n), k < n/2 */
    int *rev;        /* bit reversal permutation */
} SUF(fftplan);

static int SUF(plan_init)(SUF(fftplan) *p, int n) {
    p->n = n;
    p->log2n = ilog2_exact(n);
    if (p->log2n < 0) return -1;
    p->tw = (SUF(cplx) *)malloc(sizeof(SUF(cplx)) * (size_t)(n / 2 + 1));
    p->rev = (int *)malloc(sizeof(int) * (size_t)n);
    if (!p->tw || !p->rev) return -1;
    for (int k = 0; k < n / 2 + 1; ++k) {
        /* exact argument reduction: angle = 2*pi*k/n, k < n */
        double a = -2.0 * M_PI * (double)k / (double)n;
        p->tw[k].re = (REAL)cos(a);
        p->tw[k].im = (REAL)sin(a);
    }
    for (int i = 0; i < n; ++i) {
        int r = 0;
        for (int b = 0; b < p->log2n; ++b)
            if (i & (1 << b)) r |= 1 << (p->log2n - 1 - b);
        p->rev[i] = r;
    }
    return 0;
}

static void SUF(plan_free)(SUF(fftplan) *p) {
    free(p->tw);
    free(p->rev);
}

/* In-place forward FFT of n complex values with element stride `stride`. */
static void SUF(fft1d)(const SUF(fftplan) *p, SUF(cplx) *a, size_t stride) {
    const int n = p->n;
    for (int i = 0; i < n; ++i) {
        int j = p->rev[i];
        if (j > i) {
            SUF(cplx) t = a[(size_t)i * stride];
            a[(size_t)i * stride] = a[(size_t)j * stride];
            a[(size_t)j * stride] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        int step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; ++k) {
                SUF(cplx) w = p->tw[k * step];
                SUF(cplx) *u = &a[(size_t)(i + k) * stride];
                SUF(cplx) *v = &a[(size_t)(i + k + half) * stride];
                REAL tr = v->re * w.re - v->im * w.im;
                REAL ti = v->re * w.im + v->im * w.re;
                v->re = u->re - tr;
                v->im = u->im - ti;
                u->re = u->re + tr;
                u->im = u->im + ti;
            }
        }
    }
}

/*
 * Unnormalised forward 2-D real-to-complex DFT of an s x s tile, the
 * transform cuFFT performs at src/DDM.cu:189-192 with the plan of
 * src/DDM.cu:654-677 (n = {s, s}, onembed = {s, s/2+1}).
 * out[ky][kx] (ky < s, kx <= s/2), interleaved re/im, row stride s/2+1.
 * work: s*s complex scratch.
 */
static void SUF(rfft2)(const SUF(fftplan) *p, const REAL *in, SUF(cplx) *out,
                       SUF(cplx) *work) {
    const int s = p->n, h = s / 2 + 1;
    /* rows (along x), real input carried as complex with zero imaginary */
    for (int y = 0; y < s; ++y) {
        SUF(cplx) *row = work + (size_t)y * s;
        for (int x = 0; x < s; ++x) {
            row[x].re = in[(size_t)y * s + x];
            row[x].im = 0;
        }
        SUF(fft1d)(p, row, 1);
    }
    /* keep the half spectrum kx <= s/2 */
    for (int y = 0; y < s; ++y)
        for (int kx = 0; kx < h; ++kx) out[(size_t)y * h + kx] = work[(size_t)y * s + kx];
    /* columns (along y), eight at a time through contiguous copies in `work`
     * (the same transforms as in place with stride h; only faster) */
    for (int k0 = 0; k0 < h; k0 += 8) {
        const int nb = h - k0 < 8 ? h - k0 : 8;
        for (int y = 0; y < s; ++y)
            for (int b = 0; b < nb; ++b) work[(size_t)b * s + y] = out[(size_t)y * h + k0 + b];
        for (int b = 0; b < nb; ++b) SUF(fft1d)(p, work + (size_t)b * s, 1);
        for (int y = 0; y < s; ++y)
            for (int b = 0; b < nb; ++b) out[(size_t)y * h + k0 + b] = work[(size_t)b * s + y];
    }
}

/* ---------------------------------------------------------------------------
 * Simple fork/join over a range with pthreads.
 * ------------------------------------------------------------------------- */
typedef void (*SUF(range_fn))(void *ctx, long lo, long hi);

typedef struct {
    SUF(range_fn) fn;
    void *ctx;
    long lo, hi;
} SUF(range_job);

static void *SUF(range_thread)(void *arg) {
    SUF(range_job) *j = (SUF(range_job) *)arg;
    if (j->lo < j->hi) j->fn(j->ctx, j->lo, j->hi);
    return NULL;
}

static void SUF(parallel_for)(int nthreads, long n, SUF(range_fn) fn, void *ctx) {
    if (nthreads <= 1 || n < 2) {
        fn(ctx, 0, n);
        return;
    }
    if (nthreads > 256) nthreads = 256;
    if (nthreads > n) nthreads = (int)n;
    pthread_t th[256];
    SUF(range_job) jobs[256];
    long per = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].fn = fn;
        jobs[t].ctx = ctx;
        jobs[t].lo = (long)t * per;
        jobs[t].hi = jobs[t].lo + per < n ? jobs[t].lo + per : n;
        if (pthread_create(&th[t], NULL, SUF(range_thread), &jobs[t]) != 0) {
            /* could not spawn: run inline */
            SUF(range_thread)(&jobs[t]);
            th[t] = 0;
        }
    }
    for (int t = 0; t < nthreads; ++t)
        if (th[t]) pthread_join(th[t], NULL);
}

/* ---------------------------------------------------------------------------
 * Per-tile pipeline: tile cut -> FFT of every frame -> lag sums.
 * ------------------------------------------------------------------------- */
typedef struct {
    const ddmo_params *prm;
    const uint8_t *frames;
    int M, s, tile;
    size_t E;               /* s * (s/2+1) */
    SUF(cplx) *spec;        /* [N][E] */
    const SUF(fftplan) *plan;
    const int *lags;        /* sorted ascending */
    int L;
    REAL *acc;              /* [L][E] */
} SUF(tile_ctx);

static void SUF(fft_frames)(void *vctx, long lo, long hi) {
    SUF(tile_ctx) *c = (SUF(tile_ctx) *)vctx;
    const int s = c->s;
    REAL *tile = (REAL *)malloc(sizeof(REAL) * (size_t)s * s);
    SUF(cplx) *work = (SUF(cplx) *)malloc(sizeof(SUF(cplx)) * (size_t)s * s);
    double *dtile = (double *)malloc(sizeof(double) * (size_t)s * s);
    for (long f = lo; f < hi; ++f) {
        ddmo_tile(c->prm, c->frames, c->M, s, c->tile, (int)f, dtile);
        for (size_t i = 0; i < (size_t)s * s; ++i) tile[i] = (REAL)dtile[i];
        SUF(rfft2)(c->plan, tile, c->spec + (size_t)f * c->E, work);
    }
    free(tile);
    free(work);
    free(dtile);
}

/* Lag sums over element blocks: acc[l][e] = sum_{t, t+tau < N} |F_t - F_{t+tau}|^2.
 * Follows processFFT (src/DDM_kernel.cuh:85-101) without the per-term 1/s^2
 * factor (a power of two, applied exactly at the end), over the valid pair
 * set {t >= 0, t + tau <= N-1}. */
static void SUF(lag_block)(void *vctx, long lo, long hi) {
    SUF(tile_ctx) *c = (SUF(tile_ctx) *)vctx;
    const long N = c->prm->n_frames;
    const size_t E = c->E;
    const long BLK = 64;
    for (long b0 = lo * BLK; b0 < hi * BLK && b0 < (long)E; b0 += BLK) {
        long b1 = b0 + BLK < (long)E ? b0 + BLK : (long)E;
        for (int l = 0; l < c->L; ++l) {
            const long tau = c->lags[l];
            REAL *acc = c->acc + (size_t)l * E;
            for (long t = 0; t + tau < N; ++t) {
                const SUF(cplx) *A = c->spec + (size_t)t * E;
                const SUF(cplx) *B = c->spec + (size_t)(t + tau) * E;
                for (long e = b0; e < b1; ++e) {
                    REAL dx = A[e].re - B[e].re;
                    REAL dy = A[e].im - B[e].im;
                    acc[e] += dx * dx + dy * dy;
                }
            }
        }
    }
}

/*
 * Full pipeline.  iqt layout: [scale (descending)][tile][q (lambda ascending)][lag (ascending)].
 * acc_out (optional): if not NULL and acc_scale >= 0, receives the lag sums of
 * that scale (index into the descending scale list) as [lag][tile][ky][kx],
 * normalised like the reference accumulator (multiplied by 1/s^4).
 * Returns 0 or a negative error code.
 */
int SUF(ddmo_run)(const ddmo_params *prm, const uint8_t *frames, double *iqt,
                  int acc_scale, double *acc_out) {
    int rc = ddmo_check(prm);
    if (rc) return rc;
    int *scales = NULL, *lags = NULL;
    float *lambdas = NULL;
    rc = ddmo_sorted_params(prm, &scales, &lags, &lambdas);
    if (rc) return rc;
    const int S = prm->scale_count, L = prm->lag_count, Q = prm->lambda_count;
    const int M = scales[0];
    const long N = prm->n_frames;
    size_t out_off = 0;
    for (int si = 0; si < S; ++si) {
        const int s = scales[si];
        const int tps = M / s, T = tps * tps;
        const size_t E = (size_t)s * (size_t)(s / 2 + 1);
        if (prm->only_scale >= 0 && prm->only_scale != si) {
            out_off += (size_t)T * Q * L;
            continue;
        }
        const int T_run = (prm->max_tiles > 0 && prm->max_tiles < T) ? prm->max_tiles : T;
        SUF(fftplan) plan;
        if (SUF(plan_init)(&plan, s)) { rc = DDMO_ENOMEM; goto done; }
        uint8_t *mask = (uint8_t *)malloc((size_t)Q * E);
        int *counts = (int *)malloc(sizeof(int) * (size_t)Q);
        SUF(cplx) *spec = (SUF(cplx) *)malloc(sizeof(SUF(cplx)) * (size_t)N * E);
        REAL *acc = (REAL *)malloc(sizeof(REAL) * (size_t)L * E);
        if (!mask || !counts || !spec || !acc) {
            free(mask); free(counts); free(spec); free(acc);
            SUF(plan_free)(&plan);
            rc = DDMO_ENOMEM;
            goto done;
        }
        for (int q = 0; q < Q; ++q)
            counts[q] = ddmo_mask(s, lambdas[q], prm->tolerance, mask + (size_t)q * E);
        const double n2 = 1.0 / ((double)s * s) / ((double)s * s);
        for (int tile = 0; tile < T_run; ++tile) {
            if (prm->tile_list_count > 0) {
                int listed = 0;
                for (int k = 0; k < prm->tile_list_count; ++k) listed |= prm->tile_list[k] == tile;
                if (!listed) continue;
            }
            SUF(tile_ctx) c;
            c.prm = prm; c.frames = frames; c.M = M; c.s = s; c.tile = tile; c.E = E;
            c.spec = spec; c.plan = &plan; c.lags = lags; c.L = L; c.acc = acc;
            SUF(parallel_for)(prm->nthreads, N, SUF(fft_frames), &c);
            memset(acc, 0, sizeof(REAL) * (size_t)L * E);
            SUF(parallel_for)(prm->nthreads, (long)((E + 63) / 64), SUF(lag_block), &c);
            if (acc_out && si == acc_scale) {
                for (int l = 0; l < L; ++l)
                    for (size_t e = 0; e < E; ++e)
                        acc_out[((size_t)l * T + tile) * E + e] = (double)acc[(size_t)l * E + e] * n2;
            }
            /* masked azimuthal average, src/azimuthal_average.cu:316-339 */
            for (int q = 0; q < Q; ++q) {
                const uint8_t *mq = mask + (size_t)q * E;
                for (int l = 0; l < L; ++l) {
                    double v = 0.0;
                    if (counts[q] > 0) {
                        const REAL *a = acc + (size_t)l * E;
                        double sum = 0.0;
                        for (size_t e = 0; e < E; ++e)
                            if (mq[e]) sum += (double)a[e];
                        double norm = ddmo_norm(prm, lags[l]);
                        v = 2.0 * sum * n2 / (double)counts[q] * norm;
                    }
                    iqt[out_off + ((size_t)tile * Q + q) * L + l] = v;
                }
            }
        }
        out_off += (size_t)T * Q * L;
        free(mask); free(counts); free(spec); free(acc);
        SUF(plan_free)(&plan);
    }
done:
    free(scales); free(lags); free(lambdas);
    return rc;
}

/* Standalone rfft2 of one s x s tile (double in, interleaved out). */
int SUF(ddmo_rfft2)(const double *in, int s, double *out) {
    SUF(fftplan) plan;
    if (SUF(plan_init)(&plan, s)) return DDMO_EINVAL;
    REAL *t = (REAL *)malloc(sizeof(REAL) * (size_t)s * s);
    SUF(cplx) *o = (SUF(cplx) *)malloc(sizeof(SUF(cplx)) * (size_t)s * (s / 2 + 1));
    SUF(cplx) *w = (SUF(cplx) *)malloc(sizeof(SUF(cplx)) * (size_t)s * s);
    if (!t || !o || !w) { free(t); free(o); free(w); SUF(plan_free)(&plan); return DDMO_ENOMEM; }
    for (size_t i = 0; i < (size_t)s * s; ++i) t[i] = (REAL)in[i];
    SUF(rfft2)(&plan, t, o, w);
    for (size_t i = 0; i < (size_t)s * (s / 2 + 1); ++i) {
        out[2 * i] = (double)o[i].re;
        out[2 * i + 1] = (double)o[i].im;
    }
    free(t); free(o); free(w);
    SUF(plan_free)(&plan);
    return 0;
}
