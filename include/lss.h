/*
 * lss.h -- C ABI of the MI355X online template matcher.
 *
 * The drop-in boundary for the matcher of the reference's OnlineSpikesV2
 * (src/SpikeSorter/OnlineSpikesV2.h:29-56, src/SpikeSorter/OnlineSpikesV2.cu:248-1001).
 * Plain C types only: no HIP or torch types cross this interface, no C++
 * exception escapes it.  Every function returns LSS_OK (0) or a negative
 * status; lss_last_error() gives the message of the last failure on the
 * calling thread.
 *
 * One handle per probe stream and host thread; handles share nothing, so
 * several handles may run on several threads without locks.  The caller owns
 * every host buffer passed in; the library owns the device buffers.
 */
#ifndef LSS_H
#define LSS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSS_OK 0
#define LSS_ERR_ARG -1       /* bad argument or size beyond what the handle was created for */
#define LSS_ERR_IO -2        /* oss_input file missing or malformed */
#define LSS_ERR_HIP -3       /* HIP runtime failure */
#define LSS_ERR_CAPACITY -4  /* more events than the caller's buffer holds: the buffer holds the first
                                ev_cap of them, *n_ev = total, and the handle's state is as before
                                the call (repeat it with a larger buffer) */
#define LSS_ERR_STATE -5     /* call out of order (e.g. lss_flush before any chunk) */
#define LSS_ERR_TIMEOUT -6   /* a wait inside the peel kernel timed out: its workgroups were not all
                                resident at once (e.g. a group graph on a device other work holds) */
#define LSS_ERR_LIMIT -7     /* a per-batch limit of the handle was exceeded (max_events, or the peel's
                                per-column-block spike lists); not about the caller's buffer */

/* Peak-acceptance rules.  KILOSORT follows run_matching
 * (KS/template_matching.py:159-198: Cmax > Th^2, exact peak, one template per
 * sample, 100 iterations).  LSS follows the CUDA matcher
 * (myGPUhelpers.cu:419-478, OnlineSpikesV2.cu:903-1001: Cmax >= Th^2, 1e-4
 * tie tolerance per template, Th truncated to an integer, 50 iterations). */
#define LSS_SEM_KILOSORT 0
#define LSS_SEM_LSS 1

#define LSS_PREC_FP32 0          /* FP32 projection (FP32 MFMA) */
#define LSS_PREC_BF16_RESCORE 1  /* BF16 MFMA projection, FP32 accumulate, near-threshold FP32 rescore */
#define LSS_PREC_BF16X3 2        /* split-BF16 MFMA projection: each FP32 operand as a BF16 hi + lo
                                    pair, three BF16 products (lo*hi + hi*lo + hi*hi) accumulated in
                                    FP32; within 3 * 2^-16 * sum|U||Bpc| of the exact scores, so every
                                    score is decided at FP32-equivalent accuracy without rescoring */

typedef struct lss_config {
    int32_t device;        /* HIP device ordinal (OnlineSpikesV2.cu:356-357) */
    int32_t chunk;         /* samples per streaming chunk, for lss_process_chunk (default 3000) */
    int32_t max_batch;     /* largest batch, margins included (default chunk + 2*M) */
    int32_t max_iter;      /* peel iterations; 0 = semantics default (100 / 50) */
    int32_t semantics;     /* LSS_SEM_* */
    int32_t precision;     /* LSS_PREC_* */
    int32_t dedup;         /* per-template duplicate suppression within dt samples
                              (saveSpikes, OnlineSpikesV2.cu:1192-1194); -1 = on for LSS only */
    int32_t max_events;    /* event capacity per batch on the device (default 65536) */
    int32_t keep_stages;   /* nonzero: keep per-stage device copies for lss_debug_stage */
    int32_t sparse;        /* 0: dense.  1: skip exact-zero template/channel blocks and exact-zero
                              ctc rows (bit-identical to dense).  2 (default): also treat as zero a
                              projection weight at most 2^-40 of its template's largest and a ctc row
                              of s whose entries are all at most 2^-40 sqrt(nm[s] nm[t]) (the far
                              tails of spatially local templates): a score moves by under 2^-40 of
                              its scale, far below one FP32 rounding of a score near threshold */
    int32_t profile;       /* 1: time each device stage with HIP events (lss_stage_times)
                              and run the peel kernel's phase clock (lss_debug_stage
                              stage 6; it costs the peel ~5%); 2: the peel stage only
                              (two events per batch), no phase clock.  Groups: nonzero
                              times the stages in front of the peel and the peel */
    int32_t peel_split;    /* template subsets per peel column block: 0 = automatic,
                              1 = none, n > 1 = n when that geometry fits, else at
                              most n (Kilosort semantics only) */
    int32_t relabel;       /* nonzero: an event's template_id is the final cluster nearest to
                              the spike's position (cluster_centroids.npy), as saveSpikes
                              reports it (OnlineSpikesV2.cu:1186-1189); 0 (default): the
                              matched template's file index */
    int32_t preprocess;    /* LSS_PRE_* bits: preprocessing in front of whitening, per batch
                              (OnlineSpikesV2.cu:771-794); 0 (default): the samples are
                              whitened as given */
    int32_t reserved[2];
} lss_config;

/* Preprocessing steps (lss_config.preprocess), applied in this order to each
 * batch of W samples of the C mapped channels:
 *   LSS_PRE_MEAN      subtract each channel's mean over the batch
 *                     (meanRemove, OnlineSpikesV2.cu:774-779; myGPUhelpers.cu:99-119)
 *   LSS_PRE_CAR       subtract, per sample, the lower median over channels
 *                     (medianRemove, :782-786; myGPUhelpers.cu:480-543; torch.median)
 *   LSS_PRE_HIGHPASS  the oss_input's hp_filter.npy impulse response cropped to W,
 *                     applied circularly and shifted by W/2 (highpassFilter,
 *                     :1006-1056; Kilosort fft_highpass, KS/preprocessing.py:139-154):
 *                     the shift follows torch.fft.fftshift under LSS_SEM_KILOSORT
 *                     and the reference's shift_kernel under LSS_SEM_LSS (they
 *                     differ by one sample for odd W).  Under LSS_SEM_LSS a batch
 *                     longer than the filter is an error, as there (:1010-1013);
 *                     Kilosort semantics zero-pad the filter instead.
 * With any step on, lss_info.p2p is the peak-to-peak of the raw batch
 * (P2P_calc, :771,1137-1146). */
#define LSS_PRE_MEAN 1
#define LSS_PRE_CAR 2
#define LSS_PRE_HIGHPASS 4
#define LSS_PRE_ALL 7

/* One spike: (sample, template, amplitude), as OnlineSpikesPayload carries them
 * (src/SpikeSorter/onlineSpikesPayload.h:8-48).  `sample` follows Kilosort's
 * convention st = column - nt//2 + nt0min + (absolute index of column 0)
 * (KS/template_matching.py:68). */
typedef struct lss_event {
    int64_t sample;
    int32_t template_id;
    float amplitude;
} lss_event;

typedef struct lss_info {
    int32_t C, T, M, K, nC;
    int32_t nt0min, dt, semantics;
    float Th;
    int32_t max_batch, chunk, max_iter;
    int64_t batches;        /* batches processed */
    int64_t events;         /* events emitted */
    int32_t last_iters;     /* peel iterations that found spikes in the last batch */
    int32_t last_spikes;    /* spikes found in the last batch */
    double last_ms;         /* wall time of the last batch, entry to events on host */
    int64_t white_kblocks;  /* 64x16 whitening operand blocks computed per 64-sample column tile */
    int64_t proj_kblocks;   /* 64x16 projection operand blocks computed per 64-sample column tile */
    int64_t proj_kblocks_dense;  /* the same without exact-zero skipping */
    int32_t precision;      /* LSS_PREC_FP32, LSS_PREC_BF16_RESCORE or LSS_PREC_BF16X3 */
    float p2p;                   /* peak-to-peak of the last batch's raw samples (preprocessing on) */
    int64_t rescored;       /* BF16 mode: (column, template) scores rescored in FP32 (running total) */
    int64_t rescore_flips;  /* BF16 mode: of those, accept/reject decisions the rescore changed */
    int32_t peel_cw;        /* last batch's peel geometry: columns per workgroup, */
    int32_t peel_rows;      /*   score rows per thread held in registers (0: in memory), */
    int32_t peel_split;     /*   template subsets per column block, */
    int32_t peel_groups;    /*   and workgroups */
    int64_t gate_tiles_scored;   /* proposal gate, running totals: score tiles computed, */
    int64_t gate_tiles;          /*   score tiles in the batches, */
    int64_t gate_fallback_blocks; /*  column blocks scored in full by the fallback, */
    int64_t gate_blocks;         /*   and column blocks */
} lss_info;

/* Fill *cfg with defaults. */
int lss_config_init(lss_config* cfg);

/* Load an oss_input directory and allocate the device state.
 * Replaces OnlineSpikesV2::OnlineSpikesV2 / initializeSorter
 * (OnlineSpikesV2.cu:248-281,322-382) and its loaders (:454-660).
 * Required: misc.txt, Wall3.npy, wPCA.npy, whiteningMat.npy, and either
 * iCC.npy + iU.npy or the geometry xc.npy + yc.npy.  The tables the launcher
 * derives from these (run_online_spikes.py:609-667) are computed at load time
 * when their files are absent: ctc on the GPU (FP64 accumulation rounded to
 * f32), iCC and iU from the geometry and Wall3; templates.npy is optional
 * (M and C then come from wPCA and Wall3). */
int lss_create(const char* oss_input_dir, const lss_config* cfg, struct lss_handle** out);

/* lss_create over a subset of the preclustered templates: keep[0..n_keep) are
 * ascending Wall3 row indices; the kept rows of Wall3, ctc and iU are used
 * and renumbered 0..n_keep-1, as in a directory written by
 * subset_oss_input_inplace (src/Python/subset_templates.py:291-347), which
 * subsets the same tensors on disk.  Final-cluster tensors (centroids) are
 * untouched, as there. */
int lss_create_subset(const char* oss_input_dir, const lss_config* cfg, const int32_t* keep, int32_t n_keep,
                      struct lss_handle** out);

/* Match one batch whose columns are already framed with margins: W = n_samples
 * columns of int16 samples, row stride `ld` (>= C) in the host buffer,
 * channel c of the matcher read from column chan_map[c] (NULL = identity).
 * Column 0 is absolute sample first_sample.  Spikes are only accepted in
 * columns [M, W-M).  Events are sorted by sample.  feat (nullable) receives
 * [ev_cap][nC][K] PC features.  Replaces kilosortMatchingPursuit
 * (OnlineSpikesV2.h:56, OnlineSpikesV2.cu:903-1001) plus computeClosestClusters'
 * feature part (OnlineSpikesV2.cu:853-875). */
int lss_process_batch(struct lss_handle* h, const int16_t* samples, int64_t n_samples, int32_t ld,
                      const int32_t* chan_map, int64_t first_sample,
                      lss_event* ev, int64_t ev_cap, int64_t* n_ev, float* feat);

/* As lss_process_batch, with `d_samples` a device pointer (HBM-resident input;
 * chan_map is a host array). */
int lss_process_batch_device(struct lss_handle* h, const int16_t* d_samples, int64_t n_samples,
                             int32_t ld, const int32_t* chan_map, int64_t first_sample,
                             lss_event* ev, int64_t ev_cap, int64_t* n_ev, float* feat);

/* lss_process_batch_device without any device allocation: the batch's device
 * buffers (whitened batch, PC projection, scores, the peel's exchange words and
 * counters, the event staging) come from one static device array of
 * LSS_STATIC_SCRATCH_BYTES per GPU (128 MiB), held from the library's first use of the
 * array on that GPU until the process ends.  No counterpart in the reference:
 * it is for callers that cannot allocate during a batch (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The handle
 * (created beforehand, which allocates) supplies the tables.  The handle's
 * per-batch needs (about 4*(C*W + 6*C*W + T*W) bytes, more in BF16 mode, plus
 * the event staging of max_events) must fit the array: otherwise LSS_ERR_ARG
 * with the sizes.  Handles with a proposal gate or keep_stages are rejected.
 * Every call on one GPU shares that GPU's array, so calls that overlap in time
 * on one GPU (from several threads) race: serialise them.  Positions
 * (lss_last_positions) are not kept by this entry point.  Results are
 * bit-identical to lss_process_batch_device. */
#define LSS_STATIC_SCRATCH_BYTES (128ll << 20)
int lss_process_batch_device_static(struct lss_handle* h, const int16_t* d_samples, int64_t n_samples, int32_t ld,
                                    const int32_t* chan_map, int64_t first_sample, lss_event* ev, int64_t ev_cap,
                                    int64_t* n_ev, float* feat);

/* Streaming: hand over the next n_samples of the stream (first_sample = its
 * absolute index; chunks must be contiguous).  The library frames the batch
 * with the previous 2*M samples (edge-replicated before the first chunk, as
 * Kilosort's first batch, KS/io.py:616-619) and emits the events of the
 * samples whose right margin is complete.  Replaces the per-batch body of
 * runSpikeSorting (OnlineSpikesV2.cu:700-845). */
int lss_process_chunk(struct lss_handle* h, const int16_t* samples, int64_t n_samples, int32_t ld,
                      const int32_t* chan_map, int64_t first_sample,
                      lss_event* ev, int64_t ev_cap, int64_t* n_ev, float* feat);

/* Emit the events held back by the right margin, padding the end of the
 * stream by repeating its last sample (KS/io.py:620-623). */
int lss_flush(struct lss_handle* h, lss_event* ev, int64_t ev_cap, int64_t* n_ev, float* feat);

int lss_query(struct lss_handle* h, lss_info* info);

/* Copy a device stage of the last batch to host memory (test support; needs
 * keep_stages).  stage: 0 Xw [C][W] f32, 1 Bpc [K][C][W] f32, 2 B before the
 * peel [T][W] f32, 3 Cfmax of the first iteration [W] f32, 4 imax [W] i32,
 * 5 template norms [T] f32, 6 peel-kernel time per phase of the last batch,
 * u64 ticks of the 100 MHz device clock: [0..8] summed over workgroups and
 * [9..17] the largest workgroup's (top-of-iteration exchange, own scan,
 * publishing the column maxima, acceptance, waiting for and reading the
 * neighbours' maxima, events and features, own subtraction, the phase-A
 * barrier (the other waves' subtraction and scan), column reduction and FP32
 * rescoring), [18] scan passes that did work, [19] workgroups, 7 the ctc table the matcher uses, file layout
 * [T][T][2M+1] f32 (as ctc.npy, or as computed at load time without it; needs
 * neither keep_stages nor a batch); with preprocessing on: 8 the preprocessed
 * batch [C][W] f32 (the whitening input), 9 channel means [C] f32, 10 per-sample
 * medians [W] f32.
 * *bytes: in = capacity, out = size written. */
int lss_debug_stage(struct lss_handle* h, int32_t stage, void* out, int64_t* bytes);

/* Device time per stage accumulated over the batches run with cfg.profile set:
 * ms[0] input copy, ms[1] preprocessing (0 without), ms[2] whitening, ms[3]
 * proposal network and gate (0 without one), ms[4] PC conv, ms[5] projection,
 * ms[6] peel loop with events and features (LSS_NSTAGES entries);
 * *batches = batches timed.
 * reset != 0 clears the accumulators after reading.  No reference counterpart:
 * the reference's per-stage Timer scopes print nothing (src/Helpers/Timer.h:12-16). */
#define LSS_NSTAGES 7
int lss_stage_times(struct lss_handle* h, double* ms, int64_t* batches, int32_t reset);

/* Spike positions (x, y) of the events the last call returned, in the same
 * order: the PC-feature mass of each of the nC nearest channels weights the
 * channel coordinates xc/yc.  Replaces compute_spike_positions_kernel
 * (OnlineSpikesV2.cu:190-245; Kilosort's compute_spike_positions,
 * KS/postprocessing.py:33-46).  Needs xc.npy / yc.npy in the oss_input
 * directory (otherwise LSS_ERR_STATE).  xy: [cap][2]; *n = events. */
int lss_last_positions(struct lss_handle* h, float* xy, int64_t cap, int64_t* n);

int lss_destroy(struct lss_handle* h);

/* Autotuning of the peel geometry for batches of n_samples columns: every
 * candidate split of the templates over workgroups (1, and the fewest
 * subsets reaching 32, 16 and 8 register rows per thread) is timed on the
 * given device batch (median wall time of `reps` runs after one warm-up);
 * the fastest is used for that batch width from then on.  Results do not
 * depend on the split (bit-identical), so tuning on live data is safe.  No
 * reference counterpart. */
typedef struct lss_tune_result {
    int32_t n;             /* candidates timed */
    int32_t chosen;        /* index of the one kept */
    int32_t split[8];      /* template subsets per column block */
    int32_t rows[8];       /* register rows per thread */
    double ms[8];          /* median wall time of one batch */
} lss_tune_result;
int lss_autotune(struct lss_handle* h, const int16_t* d_samples, int64_t n_samples, int32_t ld, int32_t reps,
                 lss_tune_result* out);

/* ---- groups: several probe streams per launch ----
 * No reference counterpart (the reference runs one independent sorter thread
 * per probe, main.cpp:77-85).  A group matches one framed batch of every
 * stream per call with a single hipGraph replay (one batched launch per
 * stage; the peel kernel is shared by all streams).  Streams may have
 * different template counts; they share K, M and cfg.  cfg->max_batch is the
 * framed batch width W used for every stream.  All streams share one peel
 * kernel instance: the register tile of the largest T, or with Kilosort rules
 * a shorter one (16, then 32 rows per thread) with each stream's templates
 * split over ceil(T / tile rows) workgroups per column block, when that grid
 * is co-resident.  cfg->peel_split: 0 = automatic, 1 = one subset per stream,
 * n > 1 = the tile that splits the largest stream n ways when it fits.
 * lss_group_query reports each stream's peel_rows, peel_split, peel_groups. */
struct lss_group;
int lss_group_create(int32_t n_streams, const char* const* oss_input_dirs, const lss_config* cfg,
                     struct lss_group** out);
/* samples[s]: stream s's batch, [W][ld[s]] int16 (device pointers when
 * device_input != 0), or NULL when stream s has no batch this call (its
 * kernels return at once; n_ev[s] = 0); first_sample[s]: absolute index of
 * its column 0.
 * Results per stream as lss_process_batch (ev[s], ev_cap[s], n_ev[s], feat[s]
 * nullable). */
int lss_group_process(struct lss_group* g, const int16_t* const* samples, const int32_t* ld,
                      int32_t device_input, const int64_t* first_sample, lss_event* const* ev,
                      const int64_t* ev_cap, int64_t* n_ev, float* const* feat);
/* The same on launch lane `lane` (0..3; lss_group_process uses lane 0).
 * Each lane has its own HIP stream and graph; lanes may be driven from
 * several host threads at once, provided no stream has a batch in two. */
int lss_group_process_lane(struct lss_group* g, int32_t lane, const int16_t* const* samples, const int32_t* ld,
                           int32_t device_input, const int64_t* first_sample, lss_event* const* ev,
                           const int64_t* ev_cap, int64_t* n_ev, float* const* feat);
int lss_group_query(struct lss_group* g, int32_t stream, lss_info* info);
/* lss_debug_stage stage 6 (the peel kernel's phase clock of the last launch)
 * for one stream of a group.  No reference counterpart. */
int lss_group_debug_stage(struct lss_group* g, int32_t stream, int32_t stage, void* out, int64_t* bytes);
/* Device time of the launches on one lane of a group created with cfg->profile
 * set (a tick is then two graph launches, the stages in front of the peel and
 * the peel, with HIP events recorded on the lane's stream around them): ms[0] the
 * stages in front of the peel (preprocessing, whitening,
 * proposal network and gate, PC conv, projection), ms[1] the peel kernel
 * (all streams' peels in one grid); *launches = launches timed.  reset != 0
 * clears the accumulators after reading.  No reference counterpart. */
int lss_group_stage_times(struct lss_group* g, int32_t lane, double* ms, int64_t* launches, int32_t reset);
/* lss_set_proposal for one stream of a group (the launch graph is rebuilt). */
int lss_group_set_proposal(struct lss_group* g, int32_t stream, struct lss_proposal* net, float lo, float hi,
                           int32_t reach);
/* Pipelined lanes (on != 0): lanes 1..3 get per-batch buffers of their own, so
 * consecutive batches of the same streams may be in flight on two lanes at once
 * (driven from two host threads): the next batch's stages in front of the peel
 * run on the CUs the current peel leaves free, and every lane's peel waits for
 * the peel enqueued before it, so one peel grid is co-resident at a time.
 * Batches must be device-resident.  Results are those of one lane taking the
 * batches in turn; the handles' host state (event counts, the last batch) is
 * updated as each call returns.  No reference counterpart: for callers with a
 * backlog of ticks (replays, offline sorting of several probes). */
int lss_group_set_pipelined(struct lss_group* g, int32_t on);
int lss_group_destroy(struct lss_group* g);

/* ---- spike-proposal network (a learned gate; no reference counterpart) ----
 * A 4-layer 1-D convolution (32 channels, kernel 9, BF16 weights and
 * activations, FP32 accumulation) run on the whitened batch per 16-channel
 * group of the probe (channels 16g-4 .. 16g+19), giving P(spike) per (group,
 * sample).  Attached to a handle, it gates which 64-template x 64-column
 * score tiles are computed: per column block, a group whose largest P within
 * the block (widened by `reach` columns) is >= hi marks the template tiles of its own and adjacent
 * groups; a block where any group lies strictly between lo and hi is scored
 * in full (the fallback).  Trained on the GPU by lss_proposal_train_step. */
struct lss_proposal;
int32_t lss_proposal_n_params(void);
/* channels: probe channels of the inputs given to _forward/_grad; max_batch:
 * their widest batch.  Weights are He-initialised from seed. */
int lss_proposal_create(int32_t device, int32_t channels, int32_t max_batch, uint64_t seed,
                        struct lss_proposal** out);
int lss_proposal_set_params(struct lss_proposal* net, const float* params, int64_t n);
int lss_proposal_get_params(struct lss_proposal* net, float* params, int64_t n);
/* xw: host whitened batch [channels][n]; prob: host [groups][n]. */
int lss_proposal_forward(struct lss_proposal* net, const float* xw, int32_t n, float* prob);
/* Loss (weighted binary cross-entropy with logits, positives weighted by
 * pos_weight, mean over groups x samples) and its gradient [n_params] for
 * labels [groups][n] (0/1 bytes). */
int lss_proposal_grad(struct lss_proposal* net, const float* xw, const uint8_t* labels, int32_t n,
                      float pos_weight, float* loss, float* grad);
/* One Adam step (beta 0.9 / 0.999) on a batch that handle h whitens on the
 * device (d_samples: framed int16 batch [n][ld] in device memory). */
int lss_proposal_train_step(struct lss_proposal* net, struct lss_handle* h, const int16_t* d_samples, int64_t n,
                            int32_t ld, const uint8_t* labels, float pos_weight, float lr, float* loss);
int lss_proposal_destroy(struct lss_proposal* net);
/* Attach (net != NULL) or detach the gate of a handle (Kilosort semantics).
 * reach: columns on each side of a 64-column block whose P counts for it
 * (0..M).  The handle keeps a reference: destroy the network after the handle. */
int lss_set_proposal(struct lss_handle* h, struct lss_proposal* net, float lo, float hi, int32_t reach);

/* Message of the last error on this thread ("" if none). */
const char* lss_last_error(void);

/* ==== Streaming around the matcher: the live loop's window policy, event
 * bookkeeping, output formats and a recording replay source (host code) ==== */

/* The window and skip policy of runSpikeSorting (OnlineSpikesV2.cu:698-745):
 * after waiting for min_window new samples past the last fetch, one of five
 * fetches is chosen from how far the stream count is ahead of processedCt. */
typedef struct lss_window_cfg {
    int64_t min_window;      /* iMinScanWindow: samples waited for per batch */
    int64_t max_window;      /* iMaxScanWindow: samples fetched at most (the socket's maximum) */
    int32_t time_behind_ms;  /* iTimeBehind: 0 = skip ahead whenever behind */
    float samp_rate;         /* fImecSamplingRate / downsampling */
    int32_t small_skip;      /* bSmallskip */
    int32_t lookback;        /* overlap kept between batches, 2M (OnlineSpikesV2.cu:461) */
} lss_window_cfg;

#define LSS_WIN_LATEST 0       /* within max_window: fetch [processedCt, latest) to row 0 (:709-717) */
#define LSS_WIN_SKIP_LATEST 1  /* behind, time_behind 0: latest max_window samples (:720-724) */
#define LSS_WIN_CATCH_UP 2     /* behind within time_behind: max_window from processedCt (:727-731) */
#define LSS_WIN_SMALL_SKIP 3   /* too far behind: max_window from the allowed count (:734-738) */
#define LSS_WIN_BIG_SKIP 4     /* too far behind: latest max_window samples (:741-745) */

typedef struct lss_window_step {
    int32_t kind;               /* LSS_WIN_* */
    int32_t skip;               /* the batch starts at row min_window of the fetch buffer (:750-757) */
    int64_t fetch_start;        /* stream samples [fetch_start, fetch_start + fetch_n) are fetched */
    int64_t fetch_n;
    int64_t buf_row;            /* ... into the fetch buffer from this row (0 or min_window) */
    int64_t batch_row;          /* the batch is buffer rows [batch_row, batch_row + batch_n) */
    int64_t batch_n;            /* currBatchNumSamples */
    int64_t latest_ct;          /* latestCt after the fetch */
    int64_t allowed_ct;         /* allowedCt = latestCt - time_behind * rate / 1000, in FP32 as there */
    int64_t save_start;         /* saveSpikes start offset: latest_ct - batch_n + 1 (:822) */
    int64_t save_end;           /* saveSpikes end offset: batch_n - min_window */
    int64_t next_processed_ct;  /* processedCt after the batch: latest_ct - lookback (:844) */
} lss_window_step;

/* Decide the next batch from processedCt and the stream's sample count (the
 * count is taken as unchanged while fetching).  Pure function. */
int lss_window_plan(const lss_window_cfg* cfg, int64_t processed_ct, int64_t stream_ct, lss_window_step* step);

/* saveSpikes (OnlineSpikesV2.cu:1173-1206) over the events of one batch matched
 * with first_sample 0 (sample = column - M/2 + nt0min): index = sample - M -
 * lookback, time = save_start + index; a template's spike within dt of its
 * last kept time is dropped, the last time is updated before the end-valid
 * test (index >= save_end drops it), and kept spikes are inserted in time
 * order after equal times.  last_spike [n_templates] carries across batches
 * (zero-initialised, as the reference's lastSpikeTime).  Events are taken in
 * the order given (the reference's detection order is not kept by the
 * device; this matcher gives them in sample order).  src (nullable) receives
 * each kept spike's index in ev. */
int lss_save_spikes(const lss_event* ev, int64_t n, int32_t M, int32_t lookback, int32_t dt, int64_t save_start,
                    int64_t save_end, int64_t* last_spike, int32_t n_templates, lss_event* out, int64_t* src,
                    int64_t cap, int64_t* n_out);

/* spikeOutput.txt lines "time,template,amplitude,y\n" as writeSpikesToFile
 * writes them (OnlineSpikesV2.cu:1161-1164; std::ostream defaults, so floats
 * print as %g): time = recording_offset + sample.  y [n] is the spike's
 * position (nullable: 0).  The reference indexes y by the sorted position
 * against detection-order positions; here each line carries its own spike's
 * y.  out NULL: *size = bytes needed. */
int lss_format_spikes(int64_t recording_offset, const lss_event* ev, const float* y, int64_t n, char* out,
                      int64_t cap, int64_t* size);

/* OnlineSpikesPayload (src/Networking/onlineSpikesPayload.h:8-48) in cereal's
 * binary archive layout: fields in declaration order, raw little-endian,
 * vectors as a uint64 length then the elements.  `long` fields take
 * long_bytes bytes: 4 as the reference's Windows (LLP64) build sends them,
 * 8 for an LP64 build. */
typedef struct lss_payload_header {
    int64_t recording_offset;
    int64_t stream_sample_ct;
    double vrms;
    float p2p;
    int64_t process_time;
    int64_t event_stream_sample_ct;
    int16_t predict_label, label, n_trials, n_correct;
    double confidence;
} lss_payload_header;

/* out NULL: *size = bytes needed. */
int lss_payload_encode(const lss_payload_header* h, const int64_t* times, const int64_t* templates,
                       const float* amplitudes, int64_t n, int32_t long_bytes, uint8_t* out, int64_t cap,
                       int64_t* size);
/* times/templates/amplitudes NULL: only the header and *n. */
int lss_payload_decode(const uint8_t* buf, int64_t size, int32_t long_bytes, lss_payload_header* h, int64_t* times,
                       int64_t* templates, float* amplitudes, int64_t cap, int64_t* n);

/* SorterParameters (src/Networking/sorterParameters.h:6-29), same layout rules;
 * getSorterParams fills it with T, C, M, W, the sampling rate, the active
 * channels' positions and the template map (OnlineSpikesV2.cu:1076-1086). */
typedef struct lss_sorter_params {
    int64_t T, C, M, N;
    float samp_rate;
} lss_sorter_params;

int lss_sorter_params_encode(const lss_sorter_params* p, const double* chanpos, int64_t n_chanpos,
                             const int32_t* neuron_indices, int64_t n_indices, int32_t long_bytes, uint8_t* out,
                             int64_t cap, int64_t* size);
int lss_sorter_params_decode(const uint8_t* buf, int64_t size, int32_t long_bytes, lss_sorter_params* p,
                             double* chanpos, int64_t chanpos_cap, int64_t* n_chanpos, int32_t* neuron_indices,
                             int64_t indices_cap, int64_t* n_indices);

/* Replay source over a recorded .ap.bin (interleaved int16, n_chan_file
 * channels per sample, channels fastest; an imec file has 385, the last being
 * sync), memory-mapped.  channels NULL: all but the last.  Fetches return
 * [n][n_channels] int16 rows as the reference's socket fetch does for its
 * channel subset (dataSocket.cpp:103-168) and verify_batch.py:33-39 reads. */
struct lss_replay;
int lss_replay_open(const char* path, int32_t n_chan_file, const int32_t* channels, int32_t n_channels,
                    struct lss_replay** out);
int lss_replay_samples(struct lss_replay* r, int64_t* n_samples, int32_t* n_channels);
/* *n_out = samples copied (fewer at the end of the file) */
int lss_replay_fetch(struct lss_replay* r, int64_t start, int64_t n, int16_t* out, int64_t* n_out);
int lss_replay_close(struct lss_replay* r);

/* ==== Long-run z-score baseline of binned spike counts (FP64 on the GPU) ====
 * The Decoder's ZScoreSdmProcessor (src/Decoder/ZScoreSdmProcessor.cpp): spikes
 * are counted per bin of bin_samples samples (bin = sample / bin_samples,
 * C++ division; onSpikes :85-94); a completed bin's count updates a running
 * mean and M2 by Welford's method while the baseline is open (:96-114) and,
 * once frozen, is scored z = (count - mean) / sd with direction +1 / -1 beyond
 * +-trigger_z (:116-140).  Series 0 is the total count (the reference's one
 * series), series 1 + t counts template t.  Counts wait in a ring of
 * ring_bins bins on the device; an event for a bin already closed is counted
 * as late and dropped, one past the ring makes lss_zscore_complete fail. */
struct lss_zscore;
/* max(1, llround(bin_ms / 1000 * rate)) (:27-28) */
int lss_zscore_bin_samples(double bin_ms, float samp_rate, int64_t* bin_samples);
int lss_zscore_create(int32_t device, int32_t n_templates, int64_t bin_samples, int32_t ring_bins,
                      struct lss_zscore** out);
/* count events (host array, any order; e.g. what a process call returned) */
int lss_zscore_add(struct lss_zscore* z, const lss_event* ev, int64_t n);
/* close every open bin that ends at or before end_sample, in order; with z_out
 * (and dir_out) each closed bin's z [n_closed][1 + n_templates] (0 while the
 * baseline is open or its sd is not positive). */
int lss_zscore_complete(struct lss_zscore* z, int64_t end_sample, float* z_out, int8_t* dir_out, int64_t cap_bins,
                        int64_t* n_closed);
/* freeze the baseline at mean/sd [1 + n_templates] (the reference takes them
 * from the operator, :36-70); NULL: at the running mean and the sample
 * standard deviation sqrt(M2 / (n - 1)). */
int lss_zscore_freeze(struct lss_zscore* z, const double* mean, const double* sd, double trigger_z);
/* running statistics [1 + n_templates] (each nullable), late events, first open bin */
int lss_zscore_stats(struct lss_zscore* z, int64_t* bins_seen, double* mean, double* m2, int64_t* late,
                     int64_t* first_open);
int lss_zscore_destroy(struct lss_zscore* z);

#ifdef __cplusplus
}
#endif

#endif /* LSS_H */
