// Spike-proposal network: a 4-layer 1-D convolution over the whitened batch,
// run per 16-channel group of the probe (weights shared across groups), whose
// output P(spike) per (group, sample) gates which 64-template x 64-column
// score tiles the matcher computes.  No counterpart in the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace lss {

constexpr int PN_GROUP = 16;                 // channels per group
constexpr int PN_HALO = 4;                   // neighbour channels read on each side
constexpr int PN_CIN = PN_GROUP + 2 * PN_HALO;
constexpr int PN_CH = 32;                    // hidden channels
constexpr int PN_KW = 9;                     // kernel width ('same' padding of 4)
constexpr int PN_PAD = PN_KW / 2;
constexpr int PN_TILE = 128;                 // output columns per convolution workgroup
constexpr int PN_SLAB = 256;                 // columns per weight-gradient workgroup

// packed parameter order (FP32 master copy): W1 [32][24][9], b1, W2 [32][32][9], b2,
// W3 [32][32][9], b3, W4 [1][32][9], b4
constexpr int PN_W1 = 0;
constexpr int PN_B1 = PN_W1 + PN_CH * PN_CIN * PN_KW;
constexpr int PN_W2 = PN_B1 + PN_CH;
constexpr int PN_B2 = PN_W2 + PN_CH * PN_CH * PN_KW;
constexpr int PN_W3 = PN_B2 + PN_CH;
constexpr int PN_B3 = PN_W3 + PN_CH * PN_CH * PN_KW;
constexpr int PN_W4 = PN_B3 + PN_CH;
constexpr int PN_B4 = PN_W4 + PN_CH * PN_KW;
constexpr int PN_NPARAMS = PN_B4 + 1;

// MFMA operand layouts of the BF16 weights: forward [co][ci*9 + k] and, for the
// data gradient, [ci][co*9 + 8 - k]; rows padded to a multiple of 16
constexpr int pn_kpad(int ci) { return (ci * PN_KW + 15) / 16 * 16; }
constexpr int PN_KP1 = pn_kpad(PN_CIN);      // 224
constexpr int PN_KP = pn_kpad(PN_CH);        // 288

struct PnWeights {            // device pointers
    float* params;            // [PN_NPARAMS] FP32 master
    __bf16* wf1; __bf16* wf2; __bf16* wf3;   // forward, [32][KP]
    __bf16* wb2; __bf16* wb3;                // data gradient of layers 2 and 3, [32][KP]
    __bf16* wt1; __bf16* wt2; __bf16* wt3;   // inference, tap-major [32][k * CI + ci] (PN_KT1 / PN_KP wide)
};

// tap-major reduction of layer 1 (9 taps x 24 channels), padded to a multiple of 16
constexpr int PN_KT1 = (PN_KW * PN_CIN + 15) / 16 * 16;   // 224

struct PnBuffers {            // per-batch activations, [G][rows][Wp]
    __bf16* a0; __bf16* a1; __bf16* a2; __bf16* a3;
    float* z; float* p;
    float* dz; __bf16* d3; __bf16* d2; __bf16* d1;
    float* part; float* bpart;               // weight-gradient partials
    float* lpart;                            // loss partials
    int G, Wp, max_W;
};

// gate: per 64-column block n and 64-template tile m, mask[n * MT + m] = 1 when scored
struct PnGate {
    uint8_t* mask;
    const int* mt_glo; const int* mt_ghi;    // group span of each template tile
    int MT;
    float lo, hi;                            // P <= lo: no spike; P >= hi: spike; between: fallback
    int reach;                               // columns each side of a block whose P counts for it
    unsigned long long* stats;               // tiles scored, tiles, fallback blocks, blocks (running)
};

constexpr int PN_MAX_JOBS = 8;   // streams per batched launch

// one stream's forward pass: input (whitened rows, data at column xoff), weights, activations
struct PnFwdJob {
    const float* xw; int64_t xs; int xoff; int C; int W;
    PnWeights wt;
    PnBuffers b;
    const int16_t* const* active;    // nullable: the job is skipped when *active is null
};
struct PnGateJob {
    PnBuffers b;
    PnGate gt;
    int W;
    const int16_t* const* active;
};
void pn_forward_batch(const PnFwdJob* jobs, int n, hipStream_t s);
// P(spike) of each job in one fused launch (input, three convolutions and the head per
// column tile with the activations in LDS); writes b.p and b.z only (no activations)
void pn_infer_batch(const PnFwdJob* jobs, int n, hipStream_t s);
void pn_gate_batch(const PnGateJob* jobs, int n, hipStream_t s);

int pn_groups(int C);
size_t pn_part_floats(int G, int Wp);

// forward from the matcher's whitened buffer (rows = channels, data at column xoff)
// active (nullable): the launches do nothing when *active is null (an idle stream of a group)
void pn_forward(const PnWeights& wt, const PnBuffers& b, const float* xw, int64_t xs, int xoff, int C, int W,
                hipStream_t s, const int16_t* const* active = nullptr);
void pn_gate(const PnBuffers& b, const PnGate& gt, int W, hipStream_t s, const int16_t* const* active = nullptr);
void pn_infer(const PnWeights& wt, const PnBuffers& b, const float* xw, int64_t xs, int xoff, int C, int W,
              hipStream_t s, const int16_t* const* active = nullptr);
// loss and gradients (into grad [PN_NPARAMS]) for labels y [G][Wp] (0/1 bytes)
void pn_backward(const PnWeights& wt, const PnBuffers& b, const uint8_t* y, int W, float pos_weight,
                 float* grad, hipStream_t s);
void pn_adam(const PnWeights& wt, const float* grad, float* m, float* v, float lr, float beta1, float beta2,
             float eps, int step, hipStream_t s);
void pn_pack(const PnWeights& wt, hipStream_t s);

}  // namespace lss
