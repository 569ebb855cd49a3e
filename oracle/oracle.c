/*
 * CPU oracle for the online template matcher -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker, never the product: only tests/, the smoke check
 * and the CPU-baseline leg of bench.py load it.  The shipped path is the HIP
 * library under livespikesorter_amd/csrc and has no CPU fallback.
 *
 * It restates, step by step, Kilosort 4.0.13's run_matching and the feature
 * extraction in extract() as vendored by the reference
 * (src/Python/run_online_spikes/Lib/site-packages/kilosort/template_matching.py,
 * abbreviated KS/template_matching.py below), plus an "LSS" switch that
 * follows the deviations of the CUDA matcher in
 * src/SpikeSorter/OnlineSpikesV2.cu:903-1001 and src/SpikeSorter/myGPUhelpers.cu.
 *
 * One source, two builds: -DREAL=double (liborc64, the FP64 oracle) and
 * -DREAL=float (liborc32, the FP32 restatement timed as the CPU baseline).
 * Exported names carry the ORC_PREFIX (orc64_ / orc32_).
 *
 * Array layouts (row-major, C order):
 *   x      int16 [W][ld]       interleaved samples, channel fastest
 *   Wr     f32   [C][C]        whitening matrix (includes the int16 gain)
 *   wPCA   f32   [K][M]        temporal PCs
 *   U      f32   [T][K][C]     spatial templates (Kilosort Wall3)
 *   ctc    f32   [T][T][2M+1]  ctc[t][s][l] (KS/template_matching.py:110-121)
 *   iCC    i64   [nC][C]       nearest channels of each channel
 *   iU     i64   [T]           main channel of each template
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <omp.h>

#define NTHR(p) ((p)->threads > 0 ? (p)->threads : omp_get_max_threads())

#ifndef REAL
#define REAL double
#endif
#ifndef ORC_PREFIX
#define ORC_PREFIX orc64_
#endif
#define ORC_CAT2(a, b) a##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define ORC(name) ORC_CAT(ORC_PREFIX, name)

typedef struct {
    int C, T, M, K, nC;
    double Th;        /* Th_learned; caller truncates it for LSS (OnlineSpikesV2.cu:485) */
    int max_iter;     /* 100 Kilosort (KS/template_matching.py:157), 50 LSS (OnlineSpikesV2.cu:933) */
    int semantics;    /* 0 Kilosort, 1 LSS */
    int threads;      /* OpenMP threads, <= 0 for the runtime default */
} orc_params;

#define SEM_KILOSORT 0
#define SEM_LSS 1

/* ---- stage 1: int16 -> real, whitening Xw[c][w] = sum_c' Wr[c][c'] x[w][map c'] ----
 * KS: X = whiten_mat @ X (src/Python/OnlineKilosort.py:137-142); LSS: cuBLAS SGEMM
 * via matMul (OnlineSpikesV2.cu:797-801, myGPUhelpers.cu:722-738). */
void ORC(whiten)(const orc_params *p, const int16_t *x, long W, long ld, const int *chan_map,
                 const float *Wr, REAL *Xw)
{
    const int C = p->C;
    REAL *xt = (REAL *)malloc(sizeof(REAL) * (size_t)C * W);
    for (long w = 0; w < W; ++w)
        for (int c = 0; c < C; ++c)
            xt[(size_t)c * W + w] = (REAL)x[w * ld + (chan_map ? chan_map[c] : c)];
    #pragma omp parallel for schedule(static) num_threads(NTHR(p))
    for (int c = 0; c < C; ++c) {
        REAL *o = Xw + (size_t)c * W;
        for (long w = 0; w < W; ++w) o[w] = 0;
        for (int c2 = 0; c2 < C; ++c2) {
            const REAL a = (REAL)Wr[(size_t)c * C + c2];
            if (a == 0) continue;
            const REAL *xi = xt + (size_t)c2 * W;
            for (long w = 0; w < W; ++w) o[w] += a * xi[w];
        }
    }
    free(xt);
}

/* ---- stage 2: 6-PC temporal convolution, 'same' zero padding, no flip ----
 * Bpc[c][k][w] = sum_m Xw[c][w + m - M/2] wPCA[k][m]
 * KS: conv1d(X.unsqueeze(1), W.unsqueeze(1), padding=nt//2) (KS/template_matching.py:145);
 * LSS: conv1d kernel (myGPUhelpers.cu:356-400). */
void ORC(pcconv)(const orc_params *p, const REAL *Xw, long W, const float *wPCA, REAL *Bpc)
{
    const int C = p->C, K = p->K, M = p->M, h = M / 2;
    #pragma omp parallel for schedule(static) num_threads(NTHR(p))
    for (int c = 0; c < C; ++c) {
        const REAL *xr = Xw + (size_t)c * W;
        for (int k = 0; k < K; ++k) {
            REAL *o = Bpc + ((size_t)c * K + k) * W;
            const float *f = wPCA + (size_t)k * M;
            for (long w = 0; w < W; ++w) {
                REAL acc = 0;
                for (int m = 0; m < M; ++m) {
                    const long wi = w + m - h;
                    if (wi >= 0 && wi < W) acc += xr[wi] * (REAL)f[m];
                }
                o[w] = acc;
            }
        }
    }
}

/* ---- stage 3: spatial projection B[t][w] = sum_{k,c} U[t,k,c] Bpc[c][k][w] ----
 * KS: einsum('ijk, kjl -> il', U, B) (KS/template_matching.py:146);
 * LSS: cross_correlation_kernel (myGPUhelpers.cu:263-337).  Exact zeros of U
 * are skipped (0*x adds nothing). */
void ORC(project)(const orc_params *p, const REAL *Bpc, long W, const float *U, REAL *B)
{
    const int C = p->C, K = p->K, T = p->T;
    #pragma omp parallel for schedule(dynamic, 4) num_threads(NTHR(p))
    for (int t = 0; t < T; ++t) {
        REAL *o = B + (size_t)t * W;
        for (long w = 0; w < W; ++w) o[w] = 0;
        for (int k = 0; k < K; ++k)
            for (int c = 0; c < C; ++c) {
                const REAL u = (REAL)U[((size_t)t * K + k) * C + c];
                if (u == 0) continue;
                const REAL *b = Bpc + ((size_t)c * K + k) * W;
                for (long w = 0; w < W; ++w) o[w] += u * b[w];
            }
    }
}

/* nm[t] = sum_{k,c} U[t,k,c]^2 (KS/template_matching.py:141; LSS: cublasSdot, OnlineSpikesV2.cu:624-633) */
void ORC(norms)(const orc_params *p, const float *U, REAL *nm)
{
    const size_t KC = (size_t)p->K * p->C;
    for (int t = 0; t < p->T; ++t) {
        REAL s = 0;
        for (size_t i = 0; i < KC; ++i) {
            const REAL u = (REAL)U[(size_t)t * KC + i];
            s += u * u;
        }
        nm[t] = s;
    }
}

/* ---- stages 4-6: the peel loop (run_matching, KS/template_matching.py:157-198) ----
 * B is updated in place; Xres (nullable) is the residual updated for features.
 * Outputs, in detection order (iteration, then sample; LSS: then template):
 * out_w (batch column), out_t, out_amp, out_iter.  Returns the spike count,
 * or -1 on capacity overflow.  *n_iter receives the number of iterations that
 * found spikes. */
/* Decision margins (margins nullable, [3][W], each the minimum over the
 * iterations run): how close the peel came, at column w, to deciding
 * differently -- [0] threshold: |Cfmax - Th^2| / Th^2 where w is (within
 * MARGIN_NEAR of) the peak of its pooling window; [1] template: (best - second
 * best Cf) / best; [2] pooling: |Cfmax[w] - max of the other columns of the
 * window| / the larger of the two.  [1] and [2] count only where the window's
 * scores reach Th^2 (1 - MARGIN_NEAR); elsewhere they stay +inf.  Test support:
 * a spike-list difference between two runs is a near tie when a margin around
 * it is below the runs' score error. */
#define MARGIN_NEAR 1e-3
long ORC(peel_m)(const orc_params *p, REAL *B, REAL *Xres, long W, const REAL *nm,
                 const float *ctc, const float *U, const float *wPCA,
                 long cap, long *out_w, int *out_t, REAL *out_amp, int *out_iter, int *n_iter,
                 REAL *first_Cfmax, int *first_imax, REAL *margins);

long ORC(peel)(const orc_params *p, REAL *B, REAL *Xres, long W, const REAL *nm,
               const float *ctc, const float *U, const float *wPCA,
               long cap, long *out_w, int *out_t, REAL *out_amp, int *out_iter, int *n_iter,
               REAL *first_Cfmax, int *first_imax)
{
    return ORC(peel_m)(p, B, Xres, W, nm, ctc, U, wPCA, cap, out_w, out_t, out_amp, out_iter, n_iter,
                       first_Cfmax, first_imax, NULL);
}

long ORC(peel_m)(const orc_params *p, REAL *B, REAL *Xres, long W, const REAL *nm,
                 const float *ctc, const float *U, const float *wPCA,
                 long cap, long *out_w, int *out_t, REAL *out_amp, int *out_iter, int *n_iter,
                 REAL *first_Cfmax, int *first_imax, REAL *margins)
{
    const int T = p->T, M = p->M, K = p->K, C = p->C, L = 2 * M + 1, h = M / 2;
    const int lss = p->semantics == SEM_LSS;
    const REAL Th2 = (REAL)(p->Th * p->Th);
    REAL *Cfmax = (REAL *)malloc(sizeof(REAL) * W);
    REAL *Cf2 = (REAL *)malloc(sizeof(REAL) * W);     /* second largest Cf of each column */
    REAL *Cmax = (REAL *)malloc(sizeof(REAL) * W);
    if (margins)
        for (long i = 0; i < 3 * W; ++i) margins[i] = INFINITY;
    int *imax = (int *)malloc(sizeof(int) * W);
    REAL *rnm = (REAL *)malloc(sizeof(REAL) * T);
    long *pw = (long *)malloc(sizeof(long) * (W + 1) * (lss ? T : 1));
    int *pt = (int *)malloc(sizeof(int) * (W + 1) * (lss ? T : 1));
    REAL *pa = (REAL *)malloc(sizeof(REAL) * (W + 1) * (lss ? T : 1));
    REAL *wave = NULL;
    if (Xres) {
        /* wave[t][m][c] = sum_k U[t,k,c] wPCA[k][m] (einsum 'ijk, jl -> kil', :197) */
        wave = (REAL *)malloc(sizeof(REAL) * (size_t)T * M * C);
        for (int t = 0; t < T; ++t)
            for (int m = 0; m < M; ++m)
                for (int c = 0; c < C; ++c) {
                    REAL s = 0;
                    for (int k = 0; k < K; ++k)
                        s += (REAL)U[((size_t)t * K + k) * C + c] * (REAL)wPCA[(size_t)k * M + m];
                    wave[((size_t)t * M + m) * C + c] = s;
                }
    }
    for (int t = 0; t < T; ++t) rnm[t] = (REAL)1 / nm[t];   /* LSS reciprocal_kernel */
    long nsp = 0;
    int it_found = 0;
    for (int it = 0; it < p->max_iter; ++it) {
        /* Cf = relu(B)^2 / nm, zeroed in the first and last nt columns (:159-166);
         * Cfmax, imax = max over templates, first index on ties (:167) */
        #pragma omp parallel for schedule(static) num_threads(NTHR(p))
        for (long w = 0; w < W; ++w) {
            REAL best = lss ? -FLT_MAX : -INFINITY, second = -INFINITY;
            int bi = 0;
            const int edge = (w < M || w >= W - M);
            for (int t = 0; t < T; ++t) {
                REAL v = 0;
                if (!edge) {
                    REAL b = B[(size_t)t * W + w];
                    b = b > 0 ? b : 0;
                    v = lss ? (b * b) * rnm[t] : (b * b) / nm[t];
                }
                if (v > best) { second = best; best = v; bi = t; }
                else if (v > second) second = v;
            }
            Cfmax[w] = best;
            Cf2[w] = second;
            imax[w] = bi;
        }
        if (it == 0 && first_Cfmax) {
            memcpy(first_Cfmax, Cfmax, sizeof(REAL) * W);
            for (long w = 0; w < W; ++w) first_imax[w] = imax[w];
        }
        /* Cmax = max_pool1d(Cfmax, 2nt+1, stride 1, padding nt) (:168) */
        for (long w = 0; w < W; ++w) {
            REAL m = -INFINITY;
            const long lo = w - M < 0 ? 0 : w - M, hi = w + M >= W ? W - 1 : w + M;
            for (long j = lo; j <= hi; ++j) if (Cfmax[j] > m) m = Cfmax[j];
            Cmax[w] = m;
        }
        if (margins) {
            #pragma omp parallel for schedule(static) num_threads(NTHR(p))
            for (long w = M; w < W - M; ++w) {
                const REAL cf = Cfmax[w], cm = Cmax[w];
                REAL other = -INFINITY;     /* largest Cfmax of the window's other columns */
                const long lo = w - M < 0 ? 0 : w - M, hi = w + M >= W ? W - 1 : w + M;
                for (long j = lo; j <= hi; ++j) if (j != w && Cfmax[j] > other) other = Cfmax[j];
                REAL *mg = margins;
                if (cm > 0 && (cm - cf) <= (REAL)MARGIN_NEAR * cm) {
                    REAL m0 = cf - Th2;
                    m0 = (m0 < 0 ? -m0 : m0) / Th2;
                    if (m0 < mg[w]) mg[w] = m0;
                }
                if (cm >= Th2 * (REAL)(1 - MARGIN_NEAR)) {
                    if (cf > 0) {
                        const REAL m1 = (cf - Cf2[w]) / cf;
                        if (m1 < mg[W + w]) mg[W + w] = m1;
                    }
                    const REAL big = cf > other ? cf : other;
                    if (big > 0) {
                        REAL m2 = cf - other;
                        m2 = (m2 < 0 ? -m2 : m2) / big;
                        if (m2 < mg[2 * W + w]) mg[2 * W + w] = m2;
                    }
                }
            }
        }
        /* acceptance */
        long n = 0;
        if (!lss) {
            /* Cmax > Th^2 and |Cmax - Cfmax| < 1e-9, template imax (:172-181) */
            for (long w = 0; w < W; ++w) {
                const REAL cm = Cmax[w];
                REAL d = cm - Cfmax[w];
                if (d < 0) d = -d;
                if (cm > Th2 && d < (REAL)1e-9) {
                    pw[n] = w; pt[n] = imax[w];
                    pa[n] = B[(size_t)imax[w] * W + w] / nm[imax[w]];   /* :188 */
                    ++n;
                }
            }
        } else {
            /* every (t, w) with |Cmax[w] - Cf[t][w]| < 1e-4 and Cmax >= Th^2
             * (find_matching_indices_kernel, myGPUhelpers.cu:419-478), listed by (w, t) */
            for (long w = 0; w < W; ++w) {
                const REAL cm = Cmax[w];
                if (!(cm >= Th2)) continue;
                if (w < M || w >= W - M) continue;   /* Cf is 0 there and Th^2 > 0 */
                for (int t = 0; t < T; ++t) {
                    REAL b = B[(size_t)t * W + w];
                    b = b > 0 ? b : 0;
                    const REAL cf = (b * b) * rnm[t];
                    REAL d = cm - cf;
                    if (d < 0) d = -d;
                    if (d < (REAL)1e-4) {
                        pw[n] = w; pt[n] = t;
                        pa[n] = B[(size_t)t * W + w] * rnm[t];   /* compute_amps_kernel */
                        ++n;
                    }
                }
            }
        }
        if (n == 0) break;
        it_found = it + 1;
        if (nsp + n > cap) { nsp = -1; break; }
        for (long i = 0; i < n; ++i) {
            out_w[nsp + i] = pw[i]; out_t[nsp + i] = pt[i];
            out_amp[nsp + i] = pa[i]; out_iter[nsp + i] = it;
        }
        nsp += n;
        /* subtraction in two interleaved sub-passes (n = 2, :195-198); spikes of one
         * sub-pass are applied one after another */
        const int npass = lss ? 1 : 2;
        for (int j = 0; j < npass; ++j) {
            for (long i = j; i < n; i += npass) {
                const long w0 = pw[i];
                const int s = pt[i];
                const REAL a = pa[i];
                #pragma omp parallel for schedule(static) num_threads(NTHR(p))
                for (int t = 0; t < T; ++t) {
                    REAL *brow = B + (size_t)t * W;
                    const float *cr = ctc + ((size_t)t * T + s) * L;
                    for (int l = 0; l < L; ++l) {
                        const long w = w0 + l - M;
                        if (w >= 0 && w < W) brow[w] -= a * (REAL)cr[l];
                    }
                }
                if (Xres) {
                    for (int m = 0; m < M; ++m) {
                        const long w = w0 + m - h;
                        if (w < 0 || w >= W) continue;
                        const REAL *wv = wave + ((size_t)s * M + m) * C;
                        for (int c = 0; c < C; ++c) Xres[(size_t)c * W + w] -= a * wv[c];
                    }
                }
            }
        }
    }
    *n_iter = it_found;
    free(Cfmax); free(Cf2); free(Cmax); free(imax); free(rnm); free(pw); free(pt); free(pa); free(wave);
    return nsp;
}

/* ---- PC features on the final residual (KS/template_matching.py:50-51) ----
 * feat[i][j][k] = sum_m Xres[iCC[j][iU[t_i]]][w_i + m - M/2] wPCA[k][m]
 *               + amp_i * U[t_i, k, iCC[j][iU[t_i]]]          (Ucc, :25) */
void ORC(features)(const orc_params *p, const REAL *Xres, long W, const float *wPCA, const float *U,
                   const int64_t *iCC, const int64_t *iU, long n, const long *sw, const int *st,
                   const REAL *amp, REAL *feat)
{
    const int C = p->C, K = p->K, M = p->M, nC = p->nC, h = M / 2;
    #pragma omp parallel for schedule(static) num_threads(NTHR(p))
    for (long i = 0; i < n; ++i) {
        const int t = st[i];
        for (int j = 0; j < nC; ++j) {
            const long ch = iCC[(size_t)j * C + iU[t]];
            for (int k = 0; k < K; ++k) {
                REAL s = 0;
                for (int m = 0; m < M; ++m) {
                    const long w = sw[i] + m - h;
                    if (w >= 0 && w < W) s += Xres[(size_t)ch * W + w] * (REAL)wPCA[(size_t)k * M + m];
                }
                s += amp[i] * (REAL)U[((size_t)t * K + k) * C + ch];
                feat[((size_t)i * nC + j) * K + k] = s;
            }
        }
    }
}

/* Whole batch: whitening, conv, projection, peel, features.  Spikes are
 * returned sorted by (column, detection order), the order extract() returns
 * after its final argsort (KS/template_matching.py:80-83) with ties kept in
 * detection order.  Stage buffers are optional (NULL to skip). */
static long ORC(match_whitened)(const orc_params *p, REAL *Xw, long W, const float *wPCA, const float *U,
                                const float *ctc, const int64_t *iCC, const int64_t *iU,
                                long cap, long *out_w, int *out_t, REAL *out_amp, int *out_iter, REAL *out_feat,
                                int *n_iter, REAL *stage_Xw, REAL *stage_Bpc, REAL *stage_B,
                                REAL *stage_Cfmax, int *stage_imax, REAL *margins);

long ORC(run_batch)(const orc_params *p, const int16_t *x, long W, long ld, const int *chan_map,
                    const float *Wr, const float *wPCA, const float *U, const float *ctc,
                    const int64_t *iCC, const int64_t *iU,
                    long cap, long *out_w, int *out_t, REAL *out_amp, int *out_iter, REAL *out_feat,
                    int *n_iter, REAL *stage_Xw, REAL *stage_Bpc, REAL *stage_B,
                    REAL *stage_Cfmax, int *stage_imax, REAL *margins)
{
    REAL *Xw = (REAL *)malloc(sizeof(REAL) * (size_t)p->C * W);
    ORC(whiten)(p, x, W, ld, chan_map, Wr, Xw);
    return ORC(match_whitened)(p, Xw, W, wPCA, U, ctc, iCC, iU, cap, out_w, out_t, out_amp, out_iter, out_feat,
                               n_iter, stage_Xw, stage_Bpc, stage_B, stage_Cfmax, stage_imax, margins);
}

/* The same batch from a preprocessed input xf [C][W] (FP64, channel-gathered):
 * the mean, CAR and high-pass stages in front of whitening are applied by the
 * caller (oracle/preprocess.py). */
long ORC(run_batch_f)(const orc_params *p, const double *xf, long W, const float *Wr, const float *wPCA,
                      const float *U, const float *ctc, const int64_t *iCC, const int64_t *iU,
                      long cap, long *out_w, int *out_t, REAL *out_amp, int *out_iter, REAL *out_feat,
                      int *n_iter, REAL *stage_Xw, REAL *stage_Bpc, REAL *stage_B,
                      REAL *stage_Cfmax, int *stage_imax, REAL *margins)
{
    const int C = p->C;
    REAL *Xw = (REAL *)malloc(sizeof(REAL) * (size_t)C * W);
    #pragma omp parallel for schedule(static) num_threads(NTHR(p))
    for (int c = 0; c < C; ++c) {
        REAL *o = Xw + (size_t)c * W;
        for (long w = 0; w < W; ++w) o[w] = 0;
        for (int c2 = 0; c2 < C; ++c2) {
            const REAL a = (REAL)Wr[(size_t)c * C + c2];
            if (a == 0) continue;
            const double *xi = xf + (size_t)c2 * W;
            for (long w = 0; w < W; ++w) o[w] += a * (REAL)xi[w];
        }
    }
    return ORC(match_whitened)(p, Xw, W, wPCA, U, ctc, iCC, iU, cap, out_w, out_t, out_amp, out_iter, out_feat,
                               n_iter, stage_Xw, stage_Bpc, stage_B, stage_Cfmax, stage_imax, margins);
}

/* Xw is taken over (freed here). */
static long ORC(match_whitened)(const orc_params *p, REAL *Xw, long W, const float *wPCA, const float *U,
                                const float *ctc, const int64_t *iCC, const int64_t *iU,
                                long cap, long *out_w, int *out_t, REAL *out_amp, int *out_iter, REAL *out_feat,
                                int *n_iter, REAL *stage_Xw, REAL *stage_Bpc, REAL *stage_B,
                                REAL *stage_Cfmax, int *stage_imax, REAL *margins)
{
    const int C = p->C, K = p->K, T = p->T, nC = p->nC;
    REAL *Bpc = (REAL *)malloc(sizeof(REAL) * (size_t)C * K * W);
    REAL *B = (REAL *)malloc(sizeof(REAL) * (size_t)T * W);
    REAL *nm = (REAL *)malloc(sizeof(REAL) * T);
    ORC(pcconv)(p, Xw, W, wPCA, Bpc);
    ORC(project)(p, Bpc, W, U, B);
    ORC(norms)(p, U, nm);
    if (stage_Xw) memcpy(stage_Xw, Xw, sizeof(REAL) * (size_t)C * W);
    if (stage_Bpc) memcpy(stage_Bpc, Bpc, sizeof(REAL) * (size_t)C * K * W);
    if (stage_B) memcpy(stage_B, B, sizeof(REAL) * (size_t)T * W);
    long *tw = (long *)malloc(sizeof(long) * (cap + 1));
    int *tt = (int *)malloc(sizeof(int) * (cap + 1));
    REAL *ta = (REAL *)malloc(sizeof(REAL) * (cap + 1));
    int *ti = (int *)malloc(sizeof(int) * (cap + 1));
    long n = ORC(peel_m)(p, B, Xw, W, nm, ctc, U, wPCA, cap, tw, tt, ta, ti, n_iter,
                         stage_Cfmax, stage_imax, margins);
    if (n > 0) {
        /* stable counting sort by column */
        long *cnt = (long *)calloc((size_t)W + 1, sizeof(long));
        for (long i = 0; i < n; ++i) cnt[tw[i] + 1]++;
        for (long w = 0; w < W; ++w) cnt[w + 1] += cnt[w];
        for (long i = 0; i < n; ++i) {
            const long d = cnt[tw[i]]++;
            out_w[d] = tw[i]; out_t[d] = tt[i]; out_amp[d] = ta[i]; out_iter[d] = ti[i];
        }
        free(cnt);
        if (out_feat) ORC(features)(p, Xw, W, wPCA, U, iCC, iU, n, out_w, out_t, out_amp, out_feat);
    }
    (void)nC;
    free(tw); free(tt); free(ta); free(ti);
    free(Xw); free(Bpc); free(B); free(nm);
    return n;
}
