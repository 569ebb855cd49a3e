// Internal interface between the C-ABI host layer and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/lss.h"

namespace lss {

constexpr int GEMM_BM = 64;     // template / channel rows per GEMM tile
constexpr int GEMM_BN = 64;     // sample columns per GEMM tile
constexpr int GEMM_BK = 16;     // reduction block (the unit of block-sparse skipping)
constexpr int XOFF = 32;        // zero columns left of column 0 in each whitened row
constexpr int LP = 128;         // padded lag axis of the ctc table (2M+1 <= 127)
constexpr int PEEL_THREADS = 1024;
constexpr int MIN_CW = 64;      // peel columns per workgroup: >= M, so neighbours are one step away
constexpr int MAX_CW = 512;
constexpr int MAX_NBR = 3;      // workgroups whose columns a +-M window can reach (CW >= M)
constexpr int REG_NRG = 16;     // row groups of a register-resident peel tile (64 columns)
constexpr int REG_MAX_T = 1024; // templates a register-resident tile holds (64 rows per thread)
constexpr int MAX_M = 63;
constexpr int MAX_K = 8;
constexpr int MAX_STAGE = 512;  // spikes of one iteration in a workgroup's neighbourhood
constexpr int MAX_LIST = 512;   // spikes per workgroup over a whole batch
constexpr int MAX_BAND = 512;   // FP32 rescoring: (column, template) pairs per round in a workgroup
// BF16 error bound factor: |B_bf16 - B_fp32| <= BF16_ERR * ||U_t|| * ||Bpc_bf16 over the tile's
// K-blocks||: two operand roundings of 2^-8 (product 2^-7 (1 + 2^-9)), the norm taken of the
// rounded operand (1 / (1 - 2^-8)), and FP32 accumulation over <= 4096 terms (2^-12)
constexpr float BF16_ERR = 0.0081f;

// Grid barrier words: a top counter, per-group arrival counters and per-group
// release words, each on its own 256-byte line.  Grids of at least BAR_HIER_MIN
// workgroups arrive in groups of BAR_GROUP (one arrival per group on the top
// counter) and wait on their group's release word; smaller grids use the top
// counter alone.
constexpr int BAR_LINE = 64;
constexpr int BAR_GROUP = 32;
constexpr int BAR_MAX_GROUPS = 32;
constexpr int BAR_HIER_MIN = 96;
constexpr int BAR_WORDS = BAR_LINE * (1 + 2 * BAR_MAX_GROUPS);

// Tags of the column maxima the peel publishes under Kilosort rules: the
// low 32 bits of each 64-bit word are (epoch << TAG_ITER_BITS) | (iteration +
// 1), or | TAG_DONE for the final maxima of a workgroup that has left the loop
// (valid for every later iteration).  The epoch advances once per batch; when it
// wraps the reset (done by the whitening launch) clears every word, so an old word never carries a
// current tag.
constexpr int TAG_ITER_BITS = 13;
constexpr unsigned TAG_DONE = (1u << TAG_ITER_BITS) - 1;
constexpr unsigned EPOCH_MASK = (1u << (32 - TAG_ITER_BITS)) - 1;
constexpr int MAX_ITER = 4096;
static_assert(MAX_ITER + 1 < (int)TAG_DONE, "iteration tags stay below the final tag");

// peel error flags
constexpr int ERR_BARRIER = 1;
constexpr int ERR_LIST = 2;
constexpr int ERR_SLOT = 4;
constexpr int ERR_STAGE = 8;
constexpr int ERR_BAND = 16;

struct alignas(16) Spike {   // one accepted spike inside the peel kernel
    int w;          // batch column
    int t;          // template, internal order
    float amp;
    int it;         // iteration that found it
};
static_assert(sizeof(Spike) == 16, "spikes are moved as one 16-byte word");

struct OutHeader {  // written by the peel kernel into mapped host memory
    int64_t count;
    int32_t n_iter;
    int32_t err;
    float p2p;         // written by the preprocessing (peak-to-peak of the raw batch)
    int32_t pad[3];
    int32_t pad2[4];
};
// Peel-kernel time per phase in 100 MHz realtime ticks, summed and maxed over
// workgroups: top-of-iteration exchange, phase A scan, publishing the column
// maxima (LSS rules: and the wait for the neighbours'), acceptance, the wait
// for and the reading of the neighbours' maxima (LSS rules: for their spikes),
// epilogue, phase A subtraction; then the number of phase-A passes that did
// work and the number of workgroups (summed).
constexpr int NPHASE = 9;
constexpr int PROF_WORDS = 2 * NPHASE + 2;

// word 3 of a raw buffer descriptor on gfx9-family parts (gfx950): 32-bit data format
constexpr int BUF_DWORD3 = 0x00020000;

struct PeelParams {
    float* B; int BS; int B_bytes;
    int T, W, M;
    const float* nm; const float* rnm;
    const float* ctc;                  // [T][ctc_rows][LP], ctc[s][t][l]: template t row of spike template s
    int ctc_rows;                      // row stride: T padded to the register tile height (zero rows)
    int ctc_bytes;                     // size for the buffer descriptor (register path: < 2^31)
    const int* tlo; const int* thi;    // nonzero row range of ctc[s]
    const int* orig;                   // internal -> file template index (INT_MAX past T)
    int t_first;                       // internal index of file template 0
    float Th2;
    int max_iter;
    int CW, G;
    int S, NB;                         // template subsets per column block, column blocks (G = S * NB)
    // Kilosort rules: per-subset column maxima by iteration parity, 64-bit words
    // (value << 32 | tag): scores [2][S][W], then B values, then templates (same shape)
    unsigned long long* Pa;
    const unsigned* epoch;             // this batch's epoch (advanced by the whitening launch)
    int lss;
    float* Cfmax;                      // column maxima (LSS rules)
    Spike* slots; int* slot_cnt; int slot_cap;
    Spike* list; int* list_cnt; int cap_g;
    // iter_cnt [max_iter] spike totals, followed by the peel loop's neighbour
    // synchronisation words (all zero at batch start): owners done with phase B per
    // iteration [max_iter], then per column block the iteration its owner finished
    // phase B [NB], then per workgroup the iteration it finished phase A [G]
    unsigned* bar; int* iter_cnt; int* err;
    // features (and the FP32 rescoring): Bpc element (k*CP + c, w) at (k*CP + c) * BS + w
    const float* Bpc; int CP;
    const float* U; int C, K, nC;
    const int* iCC; const int* iU;
    const float* Gt;                   // [K][K][2M-1]
    int do_feat;
    int clock;                         // the phase clock (PROF_WORDS counters) runs: the profiling instance
    // outputs
    OutHeader* hdr; lss_event* ev; float* feat; int ev_cap;
    int half; int nt0min;
    float* dbg_cfmax; int* dbg_imax;   // first-iteration copies, nullable
    unsigned long long* prof;          // PROF_WORDS counters, zeroed per batch (written when clock != 0)
    // FP32 rescoring of a BF16 projection (rescore != 0).  Every 64 x 64 score
    // tile holding a score that could reach lo2 in FP32 (its BF16 score plus the
    // BF16 error bound) has been recomputed by the FP32 projection: act[(w/64) *
    // act_mt + t/64] = 1, and those scores are the FP32 matcher's own.  The peel
    // checks, per pass, that no score of the other tiles could reach lo2; where
    // one could, the column's candidates (every template whose score lies within
    // the BF16 error bound of the column's best) are decided on FP32 scores: an
    // FP32 dot product of U2 row t with Bpc column w over the listed K-blocks of
    // t's tile, minus the BF16 value Binit[t][w], added to the current score.
    int rescore;
    float lo2;
    const uint8_t* act; int act_mt;
    const float* U2; int ldu;          // [TP][K*CP], FP32
    const int* kb_ptr; const int* kb_idx;
    const float* Binit;                // the BF16 projection before any subtraction
    // Nsq[m][w] (row stride BS): squared norm of the BF16 PC projection of column w
    // over the K-blocks listed for 64-template tile m.  |B_bf16 - B_fp32| of
    // template t is at most BF16_ERR * sqrt(nm[t] * Nsq[t / 64][w]).
    const float* Nsq;
    unsigned long long* n_rescored;    // running totals: FP32 rescores, threshold decisions flipped
    // proposal gate: score tile (64 columns, 64 templates) scored iff gate[(w/64) * gate_mt + t/64]
    const uint8_t* gate; int gate_mt;
    // optional: the stream is idle this launch when *active is null (every workgroup returns)
    const int16_t* const* active;
    // spike positions from the features (pos [ev_cap][2], nullable; channel coordinates xc, yc)
    // and, with cent (n_cent (x, y) pairs), the event's template replaced by the nearest cluster
    float* pos; const float* xc; const float* yc;
    const float* cent; int n_cent;
};

// ---- batched launch descriptors (kernel arguments) ----
constexpr int MAX_JOBS = 8;

struct WhitenJob {
    const float* Wr; const int16_t* x;
    const int16_t* const* xptr;        // when set, the input pointer is read from device memory
    const int* cmap; const int* kb_ptr; const int* kb_idx; float* Xw;
    const float* xf; int64_t xfs;      // when set, the input is the preprocessed batch xf[c][w]
    int64_t ld; int CP, W, C, XS, mtiles, ntiles;
    int cmap_id;                       // the channel map is the identity (c -> c)
    const __bf16* Wrh; const __bf16* Wrl;   // split-BF16 mode: Wr as hi and lo BF16 parts (else null)
};

// ---- preprocessing in front of whitening (lss_preprocess.hip) ----
constexpr int PRE_MEAN = 1, PRE_CAR = 2, PRE_HP = 4;
constexpr int PRE_SLAB = 128;     // samples per partial-statistics workgroup
constexpr int PRE_ROWS = 32;      // samples per CAR workgroup
constexpr int PRE_OUT = 1024;     // high-pass outputs per workgroup (256 threads x 4)
constexpr int PRE_TCH = 2048;     // taps staged in LDS at a time
constexpr int PRE_MAX_C = 1024;

struct PreJob {
    const int16_t* x; const int16_t* const* xptr;   // input [W][ld] (as the whitening job)
    int64_t ld; const int* cmap;
    int C, W, flags, nslab;
    int shift;             // high-pass: out[k] = sum_{m < ntap} X1[(k + shift + m) mod W] taps[m]
    int ntap;
    const float* taps;
    int* psum; int* pmin; int* pmax;   // [nslab][C] partial statistics
    float* mean; float* med;           // [C], [W]
    float* X1; float* Xp; int64_t PS;  // [C][PS]: after mean and CAR, after the high-pass
    float* p2p;                        // nullable (host-mapped): max - min of the raw batch
};
struct PreBatch {
    int n;
    int start_stats[MAX_JOBS + 1], start_car[MAX_JOBS + 1], start_fir[MAX_JOBS + 1];
    PreJob job[MAX_JOBS];
};
void launch_preprocess(PreBatch& b, hipStream_t s);
struct ProjectJob {
    const float* A; const float* Bm; const int* kb_ptr; const int* kb_idx; float* C;
    const __bf16* Ab; const __bf16* Bmb;  // BF16 operands (precision mode BF16); split-BF16 mode: Ab = hi
    const __bf16* Al;                     //   part of A, Al = lo part (B is split from Bm as it is staged)
    float* C2;                            // optional second copy of the result
    float* Nsq;                           // optional (BF16): per 64-row tile and column, the squared
                                          // norm of the BF16 B operand over the tile's listed K-blocks
    // BF16 mode: the BF16 projection marks act[nt * mtiles + mt] = 1 for every tile
    // holding a score whose BF16 value plus the BF16 error bound reaches lo2
    // (rnm: the templates' reciprocal norms); the FP32 projection then recomputes
    // just those tiles, counting n_rescored[0] += scores, n_rescored[1] += threshold
    // (Th2) decisions its FP32 scores change against the BF16 copy C2
    uint8_t* act;
    const float* rnm; float lo2, Th2;
    unsigned long long* n_rescored;
    const uint8_t* gate; int gate_mt;     // optional tile mask [ntile][gate_mt]: 0 = not computed
    const int16_t* const* active;         // optional: the job is skipped when *active is null
    int lda, ldb, ldc, mtiles, ntiles;
};
struct PcconvJob {
    const float* Xw; const float* wpca; float* Bpc;
    __bf16* Bpcb;                         // optional BF16 copy of the PC projection
    const int16_t* const* active;         // optional: the job is skipped when *active is null
    int XS, M, BS, CP, wblocks;
    int C;                                // channels (rows PC_ROWS to a workgroup)
    int x3;                               // split-BF16 products on MFMA (32 channel rows to a workgroup)
    const __bf16* q;                      //   its banded PCA operand, hi rows then lo rows ([n][PCX_K])
    int ncols;                            //   and the columns it computes (the batch rounded up to 64)
};
constexpr int PC_ROWS_PER_BLOCK = 4;      // channel rows of one PC conv workgroup
constexpr int PCX_ROWS = 32;              //   of one split-BF16 PC conv workgroup
constexpr int PCX_K = 80;                 // samples of its 16-column sub-tile (M <= 63: M/2 + 48 <= 80)
// Workgroups of a PC conv job: 256-column blocks x channel-row blocks.
__host__ __device__ inline int pc_job_blocks(const PcconvJob& j) {
    const int rows = j.x3 ? PCX_ROWS : PC_ROWS_PER_BLOCK;
    return j.wblocks * ((j.C + rows - 1) / rows);
}
template <class J> struct JobBatch {
    int n;
    int start[MAX_JOBS + 1];       // first block of each job; start[n] = blocks in the launch
    // nonzero when the n jobs (n dividing 8) are interleaved over the XCDs, job j on
    // the XCDs x = j (mod n), so each XCD's L2 holds the operands of 8 / n ... 1 jobs
    // instead of all of them (GEMM launches then have n * gemm_blocks_ilv blocks, the
    // PC conv n times the largest job's blocks, and start[] is not read)
    int ilv;
    J job[MAX_JOBS];
};
using WhitenBatch = JobBatch<WhitenJob>;
using ProjectBatch = JobBatch<ProjectJob>;
using PcconvBatch = JobBatch<PcconvJob>;

// Blocks a GEMM job occupies: M-tiles x N-tiles padded to a multiple of 8 (XCD order).
inline int gemm_blocks(int mtiles, int ntiles) { return mtiles * ((ntiles + 7) / 8 * 8); }
// Interleaved launches: blocks per job when n jobs share the 8 XCDs (8 / n XCDs each).
inline int gemm_blocks_ilv(int mtiles, int ntiles, int n) {
    const int per = 8 / n;
    return mtiles * ((ntiles + per - 1) / per * per);
}


struct ResetBatch {
    int n;
    unsigned* bar[MAX_JOBS];
    int* iter_cnt[MAX_JOBS];
    int n_iter[MAX_JOBS];
    int* err[MAX_JOBS];
    unsigned long long* prof[MAX_JOBS];
    unsigned* epoch[MAX_JOBS];
    unsigned long long* ptag[MAX_JOBS];
    int64_t ptag_words[MAX_JOBS];
    const int16_t* const* active[MAX_JOBS];   // nullable: a stream idle this launch keeps its counters
};

// The whitening launch also resets the peel words of its streams (the peel of
// the same batch comes after it on the stream): rb may be null.
struct WhitenLaunch {
    WhitenBatch w;
    ResetBatch r;
};
static_assert(sizeof(WhitenLaunch) <= 4096, "whitening launch must fit the kernel-argument segment");
void launch_whiten(const WhitenBatch& b, const ResetBatch* rb, hipStream_t s);
void launch_pcconv(const PcconvBatch& b, int K, hipStream_t s);
// projection kernels of a precision mode (LSS_PREC_*): FP32 MFMA; BF16 MFMA then the FP32
// recomputation of the tiles it marked; split-BF16 (three BF16 products per FP32 product)
void launch_project(const ProjectBatch& b, int precision, hipStream_t s);

void launch_norms(const float* U, int KC, int T, float* nm, float* rnm, double* nm64, hipStream_t s);
// load-time ctc table on the GPU (FP64 accumulation): U [T][K][C] internal order,
// wtw [K*K][2M+1] FP64, span [T] = [first, last+1) nonzero channel of each template;
// ctc [T][TS][LP] zero on entry; tlo/thi = the t' range of each s' whose rows matter (thi = 0
// if none): nonzero rows, or with tau > 0 rows holding an entry above tau * sqrt(nm[s'] nm[t'])
// (nm: FP64 template norms, internal order)
void launch_ctc(const float* U, const double* wtw, const int2* span, int T, int K, int C, int M, int TS, float* ctc,
                int* tlo, int* thi, const double* nm, double tau, hipStream_t s);
// Sparse level 2: template entries and ctc rows this far below their scale are treated as
// zero (2^-40; see lss_config.sparse)
constexpr double NEGLIGIBLE = 0x1p-40;

// Rows of the score tile each peel thread keeps in registers (0: tile stays in global memory).
int peel_rows(int T, int CW);

struct PeelBatch {
    int n;
    int g_start[MAX_JOBS + 1];     // first workgroup of each stream; g_start[n] = grid size
    // nonzero: in a launch where every stream is active, XCD x runs the x-th eighth
    // of the workgroups (streams, then column blocks), so a stream's neighbouring
    // blocks share an L2; zero (or an idle stream): workgroup order
    int xcd;
    PeelParams p[MAX_JOBS];
};

static_assert(sizeof(PeelBatch) <= 4096, "peel batch must fit the kernel-argument segment");

hipError_t launch_peel(const PeelBatch& b, int R, bool cooperative, hipStream_t s);



int peel_max_blocks(int device, int R);

}  // namespace lss
