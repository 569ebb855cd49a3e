// C-ABI host layer of the MI355X matcher: oss_input loading, device state,
// batch orchestration and streaming.  Mirrors the reference's OnlineSpikesV2
// (src/SpikeSorter/OnlineSpikesV2.h:29-134, OnlineSpikesV2.cu:248-1001) without
// its per-iteration host round trips: one batch is five kernel launches on
// one stream and a single synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "lss_kernels.h"
#include "lss_proposal.h"
#include "npy.hpp"

using namespace lss;

namespace {

thread_local std::string g_last_error;

struct LssError : std::runtime_error {
    int code;
    LssError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw LssError(LSS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// float -> bfloat16 bit pattern, round to nearest even (inputs are finite)
inline uint16_t bf16_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

inline float bf16_value(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return LSS_OK;
    } catch (const LssError& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return LSS_ERR_IO;
    }
}

template <class T>
T* dalloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc(&p, n * sizeof(T)));
    HIPCHK(hipMemset(p, 0, n * sizeof(T)));
    return static_cast<T*>(p);
}

template <class T>
void upload(T* d, const std::vector<T>& h) {
    if (!h.empty()) HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

std::map<std::string, std::string> read_misc(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw LssError(LSS_ERR_IO, "cannot open " + path);
    std::map<std::string, std::string> kv;
    std::string line;
    while (std::getline(in, line)) {
        auto pos = line.find(':');
        if (pos == std::string::npos) continue;
        auto trim = [](std::string s) {
            while (!s.empty() && (s.back() == '\r' || s.back() == ' ' || s.back() == '\n')) s.pop_back();
            size_t i = 0;
            while (i < s.size() && s[i] == ' ') ++i;
            return s.substr(i);
        };
        kv[trim(line.substr(0, pos))] = trim(line.substr(pos + 1));
    }
    return kv;
}

bool file_exists(const std::string& p) {
    std::ifstream f(p);
    return f.good();
}

void expect_shape(const NpyArray& a, std::vector<int64_t> want, const std::string& name) {
    if (a.shape != want) {
        std::string got = "(", exp = "(";
        for (auto s : a.shape) got += std::to_string(s) + ",";
        for (auto s : want) exp += std::to_string(s) + ",";
        throw LssError(LSS_ERR_IO, name + ": shape " + got + ") expected " + exp + ")");
    }
}

}  // namespace

namespace {
// device activations of the proposal network for batches up to max_W columns
void alloc_pn_buffers(PnBuffers& b, int G, int max_W);
void free_pn_buffers(PnBuffers& b);
}  // namespace

struct lss_proposal {
    int dev = 0, C = 0, G = 0, max_W = 0;
    hipStream_t stream = nullptr;
    PnWeights wt{};
    float *d_m = nullptr, *d_v = nullptr, *d_grad = nullptr;
    float* d_xw = nullptr;          // host-input staging [C][max_W]
    uint8_t* d_y = nullptr;         // labels [G][Wp]
    PnBuffers buf{};
    int step = 0;
    ~lss_proposal() {
        (void)hipSetDevice(dev);
        void* ptrs[] = {wt.params, wt.wf1, wt.wf2, wt.wf3, wt.wb2, wt.wb3, wt.wt1, wt.wt2, wt.wt3, d_m, d_v, d_grad,
                        d_xw, d_y};
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
        free_pn_buffers(buf);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct lss_handle {
    lss_config cfg{};
    int dev = 0;
    hipStream_t stream = nullptr;
    int C = 0, T = 0, M = 0, K = 0, nC = 0, nt0min = 0, dt = 0;
    int CP = 0, TP = 0, KCP = 0;
    float Th = 0.f;
    int sem = 0, max_iter = 0;
    int Wmax = 0, XS = 0, BS = 0;
    int peel_cap[5] = {};    // co-resident peel workgroups for R = 0, 8, 16, 32, 64
    int ctc_rows = 0, t_first = 0;
    int min_rows = 0;        // force the padded ctc row count up to this (streams sharing one peel launch)
    int group_split = 0;     // a group's stream: its template subsets in the group's peel (0: one)
    int64_t cur_first = 0;   // absolute sample of column 0 of the last batch
    bool coop = true;        // cooperative launch of the peel kernel
    bool bf16 = false;       // BF16 projection with FP32 rescoring (LSS_PREC_BF16_RESCORE)
    bool x3 = false;         // split-BF16 projection (LSS_PREC_BF16X3)

    // static device tables (internal template order)
    float *d_Wr = nullptr, *d_U2 = nullptr, *d_U = nullptr, *d_wpca = nullptr;
    float *d_nm = nullptr, *d_rnm = nullptr, *d_ctc = nullptr, *d_Gt = nullptr;
    double* d_nm64 = nullptr;
    int *d_orig = nullptr, *d_tlo = nullptr, *d_thi = nullptr, *d_iCC = nullptr, *d_iU = nullptr;
    int *d_cmap = nullptr;
    int *d_kbw_ptr = nullptr, *d_kbw_idx = nullptr, *d_kbp_ptr = nullptr, *d_kbp_idx = nullptr;
    // per-batch device buffers
    float *d_Xw = nullptr, *d_Bpc = nullptr, *d_B = nullptr;
    float* d_Cfmax = nullptr;
    Spike *d_slots = nullptr, *d_list = nullptr;
    int *d_slot_cnt = nullptr, *d_list_cnt = nullptr, *d_iter_cnt = nullptr, *d_err = nullptr;
    unsigned* d_bar = nullptr;
    unsigned long long* d_prof = nullptr;
    // BF16 projection operands, its unsubtracted result, and the rescore counters
    uint16_t *d_U2b = nullptr, *d_Bpcb = nullptr;
    uint16_t* d_U2l = nullptr;       // split-BF16: the lo part of U2 (d_U2b holds the hi part)
    uint16_t *d_Wrh = nullptr, *d_Wrl = nullptr;   // split-BF16: the whitening matrix's hi and lo parts
    uint16_t* d_pcq = nullptr;                     //   and the PC conv's banded PCA operand
    float* d_Binit = nullptr;
    float* d_Nsq = nullptr;          // per 64-template tile and column: squared norm of the BF16 operand
    uint8_t* d_act = nullptr;        // per score tile [BS/64][TP/64]: recomputed by the FP32 projection
    unsigned long long* d_rescnt = nullptr;
    // proposal gate (lss_set_proposal): the network, this handle's activations and the tile mask
    lss_proposal* prop = nullptr;
    PnBuffers pnb{};
    PnGate gate{};
    uint8_t* d_gate = nullptr;
    int* d_mt_glo = nullptr;
    int* d_mt_ghi = nullptr;
    unsigned long long* d_gate_stats = nullptr;
    int slot_cap = 0, cap_g = 0, max_groups = 0;
    int g_cap = 0, sync_words = 0;       // peel workgroups at most; words of the iteration/flag block
    // template split of the peel: per-subset column maxima [2][max_split][Wmax]
    int max_split = 1;
    int force_split = 0;                 // autotuning: the split being timed (0: none forced)
    std::map<int, int> tuned_split;      // batch width -> split chosen by lss_autotune
    // Kilosort rules: the peel's tagged column maxima (score, B value and template
    // arrays of [2][S][W] each, back to back, for the batch's S and W) and the
    // epoch of their tags
    unsigned long long* d_ptag = nullptr;
    int64_t ptag_words = 0;
    unsigned* d_epoch = nullptr;
    // debug copies
    float *d_dbgB = nullptr, *d_dbgCf = nullptr;
    int* d_dbgIm = nullptr;
    // input staging
    int16_t* d_x = nullptr;
    size_t d_x_elems = 0;
    // mapped host output
    void* h_out = nullptr;
    OutHeader* hdr = nullptr;
    lss_event* ev = nullptr;
    float* feat = nullptr;
    OutHeader* d_hdr = nullptr;
    lss_event* d_ev = nullptr;
    float* d_feat = nullptr;
    // spike positions (xc/yc present) and nearest-cluster relabelling (cfg.relabel)
    float *pos = nullptr, *d_pos = nullptr;
    float *d_xc = nullptr, *d_yc = nullptr, *d_cent = nullptr;
    int n_cent = 0;
    std::vector<float> last_pos;
    int ev_cap = 0;

    std::vector<int> orig, cmap_default, cmap_cur;
    bool cmap_id = false;                // cmap_cur maps every channel to itself
    std::vector<int> iU_int;         // main channel of each template, internal order
    std::vector<int> keep;           // template subset (file indices), empty = all
    // preprocessing in front of whitening (cfg.preprocess)
    int pre = 0;
    float* d_hp = nullptr;           // hp_filter.npy, whole impulse response
    int hp_F = 0, hp_lo = 0, hp_hi = 0;   // its length and the span of taps that matter
    int *d_psum = nullptr, *d_pmin = nullptr, *d_pmax = nullptr;
    float *d_pmean = nullptr, *d_pmed = nullptr, *d_X1 = nullptr, *d_Xp = nullptr;
    int64_t PS = 0;
    int last_W = 0;

    // streaming state
    bool started = false, flushed = false;
    int64_t next_sample = 0;
    int32_t stream_ld = 0;
    std::vector<int16_t> hist;   // last 2*M samples, [2M][C] gathered channel order
    std::vector<int16_t> frame;  // host frame of the current batch, [W][C]
    std::vector<int64_t> last_spike;

    lss_info info{};

    // stage timing (cfg.profile)
    hipEvent_t evt[LSS_NSTAGES + 1] = {};
    double stage_ms[LSS_NSTAGES] = {};
    int64_t timed_batches = 0;
    bool copy_timed = false;

    ~lss_handle() { release(); }

    void release() {
        void* ptrs[] = {d_Wr, d_U2, d_U, d_wpca, d_nm, d_rnm, d_ctc, d_Gt, d_nm64, d_orig, d_tlo, d_thi,
                        d_iCC, d_iU, d_cmap, d_kbw_ptr, d_kbw_idx, d_kbp_ptr, d_kbp_idx, d_Xw, d_Bpc, d_B,
                        d_Cfmax, d_slots, d_list, d_slot_cnt, d_list_cnt, d_iter_cnt, d_err, d_bar,
                        d_dbgB, d_dbgCf, d_dbgIm, d_x, d_prof, d_U2b, d_U2l, d_Wrh, d_Wrl, d_pcq, d_Bpcb, d_Binit, d_Nsq, d_act, d_rescnt, d_ptag, d_epoch,
                        d_gate, d_mt_glo, d_mt_ghi, d_gate_stats, d_xc, d_yc, d_cent, d_hp, d_psum, d_pmin,
                        d_pmax, d_pmean, d_pmed, d_X1, d_Xp};
        (void)hipSetDevice(dev);
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        free_pn_buffers(pnb);
        if (h_out) (void)hipHostFree(h_out);
        for (auto& e : evt)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
        h_out = nullptr;
        stream = nullptr;
    }
};

namespace {

void alloc_pn_buffers(PnBuffers& b, int G, int max_W) {
    b.G = G;
    b.max_W = max_W;
    b.Wp = (int)round_up(max_W, PN_SLAB);
    const size_t n1 = (size_t)G * b.Wp;
    b.a0 = dalloc<__bf16>(n1 * PN_CIN);
    b.a1 = dalloc<__bf16>(n1 * PN_CH);
    b.a2 = dalloc<__bf16>(n1 * PN_CH);
    b.a3 = dalloc<__bf16>(n1 * PN_CH);
    b.z = dalloc<float>(n1);
    b.p = dalloc<float>(n1);
    b.dz = dalloc<float>(n1);
    b.d3 = dalloc<__bf16>(n1 * PN_CH);
    b.d2 = dalloc<__bf16>(n1 * PN_CH);
    b.d1 = dalloc<__bf16>(n1 * PN_CH);
    b.part = dalloc<float>(pn_part_floats(G, b.Wp));
    b.bpart = dalloc<float>((size_t)G * (b.Wp / PN_SLAB) * PN_CH);
    b.lpart = dalloc<float>((n1 + 255) / 256);
}

void free_pn_buffers(PnBuffers& b) {
    void* ptrs[] = {b.a0, b.a1, b.a2, b.a3, b.z, b.p, b.dz, b.d3, b.d2, b.d1, b.part, b.bpart, b.lpart};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    b = PnBuffers{};
}

// K-block lists for block-sparse GEMMs: for each 64-row tile of A [rows][cols]
// (row-major, cols a multiple of 16), the 16-column blocks holding any nonzero.
void build_kblocks(const std::vector<float>& A, int rows, int cols, bool sparse, std::vector<int>& ptr,
                   std::vector<int>& idx) {
    const int mt = rows / GEMM_BM, kbs = cols / GEMM_BK;
    ptr.assign(mt + 1, 0);
    idx.clear();
    for (int t = 0; t < mt; ++t) {
        for (int kb = 0; kb < kbs; ++kb) {
            bool nz = !sparse;
            for (int r = 0; r < GEMM_BM && !nz; ++r) {
                const float* a = &A[(size_t)(t * GEMM_BM + r) * cols + kb * GEMM_BK];
                for (int c = 0; c < GEMM_BK; ++c)
                    if (a[c] != 0.f) { nz = true; break; }
            }
            if (nz) idx.push_back(kb);
        }
        ptr[t + 1] = (int)idx.size();
    }
    if (idx.empty()) idx.push_back(0);
}

void load(lss_handle* h, const std::string& dir_in) {
    std::string dir = dir_in;
    if (!dir.empty() && dir.back() != '/') dir += '/';
    // misc.txt (OnlineSpikesV2.cu:465-487); Th_learned parsed as float, truncated for LSS (:485)
    auto kv = read_misc(dir + "misc.txt");
    for (const char* key : {"nt0min", "numNearestChans", "Th_learned", "duplicate_spike_bins"})
        if (!kv.count(key)) throw LssError(LSS_ERR_IO, std::string("misc.txt lacks ") + key);
    h->nt0min = std::stoi(kv["nt0min"]);
    h->nC = std::stoi(kv["numNearestChans"]);
    h->Th = std::stof(kv["Th_learned"]);
    h->dt = std::stoi(kv["duplicate_spike_bins"]);
    if (h->sem == LSS_SEM_LSS) h->Th = (float)(int)h->Th;

    // Wall3 gives T (unclu_T) and K (loadPreclusterShapes, :489-497); templates.npy gives
    // M and C (loadTemplatesShape, OnlineSpikesV2.cu:454-462).  Without templates.npy, C
    // comes from Wall3 and M from wPCA, which carry the same sizes.
    auto wall3 = npy_load(dir + "Wall3.npy");
    if (wall3.shape.size() != 3) throw LssError(LSS_ERR_IO, "Wall3.npy must be [T][K][C]");
    const int T_file = (int)wall3.shape[0];
    h->T = T_file;
    h->K = (int)wall3.shape[1];
    h->C = (int)wall3.shape[2];
    // template subset (subset_templates.py:291-347 applied at load time): the kept rows of
    // the preclustered tensors, renumbered 0..n-1 in the order given (ascending)
    std::vector<int>& keep = h->keep;
    if (!keep.empty()) {
        for (size_t i = 0; i < keep.size(); ++i)
            if (keep[i] < 0 || keep[i] >= T_file || (i && keep[i] <= keep[i - 1]))
                throw LssError(LSS_ERR_ARG, "template subset must be ascending indices below Wall3's template count");
        h->T = (int)keep.size();
    }
    auto wp_a = npy_load(dir + "wPCA.npy");
    if (wp_a.shape.size() != 2) throw LssError(LSS_ERR_IO, "wPCA.npy must be [K][M]");
    h->M = (int)wp_a.shape[1];
    if (file_exists(dir + "templates.npy")) {
        auto tpl = npy_load(dir + "templates.npy");
        if (tpl.shape.size() != 3 || tpl.shape[2] != h->C || tpl.shape[1] != h->M)
            throw LssError(LSS_ERR_IO, "templates.npy must be [T][M][C] with Wall3's C and wPCA's M");
    }
    const int C = h->C, T = h->T, M = h->M, K = h->K, nC = h->nC;
    if (M % 2 == 0 || M > MAX_M) throw LssError(LSS_ERR_IO, "template length must be odd and <= 63");
    if (K < 1 || K > MAX_K) throw LssError(LSS_ERR_IO, "PC count must be 1..8");
    if (nC < 1 || nC * K > 64 || nC > C) throw LssError(LSS_ERR_IO, "numNearestChans out of range");
    auto U = wall3.as<float>();
    if (!keep.empty()) {
        const size_t row = (size_t)K * C;
        std::vector<float> Us((size_t)T * row);
        for (int t = 0; t < T; ++t) std::memcpy(&Us[t * row], &U[(size_t)keep[t] * row], row * sizeof(float));
        U.swap(Us);
    }
    auto wr_a = npy_load(dir + "whiteningMat.npy");
    expect_shape(wr_a, {C, C}, "whiteningMat.npy");
    auto Wr = wr_a.as<double>();
    expect_shape(wp_a, {K, M}, "wPCA.npy");
    auto wpca = wp_a.as<double>();

    // drift matrix (a4'): the file holds get_drift_matrix(...).T (run_online_spikes.py:628);
    // Kilosort and LSS both apply that transpose back, X <- D (Wr X) (OnlineKilosort.py:139-140).
    // It is folded into the whitening matrix in FP64 when it is not the identity.
    if (file_exists(dir + "drift_matrix.npy")) {
        auto dm_a = npy_load(dir + "drift_matrix.npy");
        expect_shape(dm_a, {C, C}, "drift_matrix.npy");
        auto Dt = dm_a.as<double>();
        bool ident = true;
        for (int i = 0; i < C && ident; ++i)
            for (int j = 0; j < C; ++j)
                if (Dt[(size_t)i * C + j] != (i == j ? 1.0 : 0.0)) { ident = false; break; }
        if (!ident) {
            std::vector<double> R((size_t)C * C, 0.0);
            for (int i = 0; i < C; ++i)
                for (int k = 0; k < C; ++k) {
                    const double d = Dt[(size_t)k * C + i];   // D[i][k] = file[k][i]
                    if (d == 0.0) continue;
                    for (int j = 0; j < C; ++j) R[(size_t)i * C + j] += d * Wr[(size_t)k * C + j];
                }
            Wr.swap(R);
        }
    }

    // channel map (loadChannelMap, OnlineSpikesV2.cu:519-530)
    h->cmap_default.resize(C);
    std::iota(h->cmap_default.begin(), h->cmap_default.end(), 0);
    if (file_exists(dir + "channelMap.npy")) {
        auto cm = npy_load(dir + "channelMap.npy");
        if (cm.numel() != C) throw LssError(LSS_ERR_IO, "channelMap.npy must have C entries");
        h->cmap_default = cm.as<int>();
    }

    // iCC / iU (prepare_extract, KS/template_matching.py:20-26)
    std::vector<int64_t> iCC, iU;
    if (file_exists(dir + "iCC.npy") && file_exists(dir + "iU.npy")) {
        auto a = npy_load(dir + "iCC.npy");
        if (a.shape.size() != 2 || a.shape[1] != C || a.shape[0] < nC)
            throw LssError(LSS_ERR_IO, "iCC.npy must be [>=nC][C]");
        iCC = a.as<int64_t>();
        auto b = npy_load(dir + "iU.npy");
        expect_shape(b, {T_file}, "iU.npy");
        iU = b.as<int64_t>();
        if (!keep.empty()) {
            std::vector<int64_t> s2(T);
            for (int t = 0; t < T; ++t) s2[t] = iU[keep[t]];
            iU.swap(s2);
        }
    } else {
        auto xa = npy_load(dir + "xc.npy"), ya = npy_load(dir + "yc.npy");
        auto xc = xa.as<double>(), yc = ya.as<double>();
        if ((int)xc.size() != C || (int)yc.size() != C) throw LssError(LSS_ERR_IO, "xc/yc must have C entries");
        iCC.assign((size_t)nC * C, 0);
        std::vector<int> ord(C);
        std::vector<double> ds(C);
        for (int c = 0; c < C; ++c) {
            for (int r = 0; r < C; ++r) ds[r] = (xc[c] - xc[r]) * (xc[c] - xc[r]) + (yc[c] - yc[r]) * (yc[c] - yc[r]);
            std::iota(ord.begin(), ord.end(), 0);
            std::stable_sort(ord.begin(), ord.end(), [&](int a2, int b2) { return ds[a2] < ds[b2]; });
            for (int j = 0; j < nC; ++j) iCC[(size_t)j * C + c] = ord[j];
        }
        iU.assign(T, 0);
        for (int t = 0; t < T; ++t) {
            double best = -1;
            for (int c = 0; c < C; ++c) {
                double s = 0;
                for (int k = 0; k < K; ++k) {
                    const double u = U[((size_t)t * K + k) * C + c];
                    s += u * u;
                }
                if (s > best) { best = s; iU[t] = c; }
            }
        }
    }
    for (auto v : iCC)
        if (v < 0 || v >= C) throw LssError(LSS_ERR_IO, "iCC out of range");
    for (auto v : iU)
        if (v < 0 || v >= C) throw LssError(LSS_ERR_IO, "iU out of range");

    // internal template order: by main channel (then file index) so that 64-template
    // tiles of the projection touch few channel blocks; identity when sparse is off.
    h->orig.resize(T);
    std::iota(h->orig.begin(), h->orig.end(), 0);
    if (h->cfg.sparse)
        std::stable_sort(h->orig.begin(), h->orig.end(), [&](int a, int b) { return iU[a] < iU[b]; });

    h->CP = (int)round_up(C, GEMM_BM);
    h->TP = (int)round_up(T, GEMM_BM);
    h->KCP = K * h->CP;
    const int CP = h->CP, TP = h->TP, KCP = h->KCP;

    // whitening matrix, padded [CP][CP]
    std::vector<float> Wrp((size_t)CP * CP, 0.f);
    for (int i = 0; i < C; ++i)
        for (int j = 0; j < C; ++j) Wrp[(size_t)i * CP + j] = (float)Wr[(size_t)i * C + j];
    // projection operand U2[t'][k*CP + c] and internal U[t'][k][c]
    std::vector<float> U2((size_t)TP * KCP, 0.f), Ui((size_t)T * K * C);
    for (int tp = 0; tp < T; ++tp) {
        const int t = h->orig[tp];
        for (int k = 0; k < K; ++k)
            for (int c = 0; c < C; ++c) {
                const float u = U[((size_t)t * K + k) * C + c];
                U2[(size_t)tp * KCP + (size_t)k * CP + c] = u;
                Ui[((size_t)tp * K + k) * C + c] = u;
            }
    }
    // template norms (FP64, internal order): the scale of the negligible-value tests below
    std::vector<double> nmh(T, 0.0);
    for (int tp = 0; tp < T; ++tp)
        for (int i = 0; i < K * C; ++i) nmh[tp] += (double)Ui[(size_t)tp * K * C + i] * Ui[(size_t)tp * K * C + i];
    // Sparse level 2: a projection weight at most 2^-40 of its template's largest is
    // dropped (its products sum to under 2^-40 * umax * sum|Bpc|, far under the FP32
    // resolution of any score near threshold), so the K-block lists skip the far tails
    // of spatially local templates.  The internal U (features, ctc) keeps every value.
    const double tau = h->cfg.sparse >= 2 ? NEGLIGIBLE : 0.0;
    if (tau > 0.0)
        for (int tp = 0; tp < T; ++tp) {
            float* row = &U2[(size_t)tp * KCP];
            float umax = 0.f;
            for (int i = 0; i < KCP; ++i) umax = std::max(umax, std::fabs(row[i]));
            const float cut = (float)(tau * umax);
            for (int i = 0; i < KCP; ++i)
                if (std::fabs(row[i]) <= cut) row[i] = 0.f;
        }
    std::vector<int> kbw_ptr, kbw_idx, kbp_ptr, kbp_idx;
    build_kblocks(Wrp, CP, CP, h->cfg.sparse != 0, kbw_ptr, kbw_idx);
    build_kblocks(U2, TP, KCP, h->cfg.sparse != 0, kbp_ptr, kbp_idx);
    h->info.white_kblocks = kbw_ptr.back();
    h->info.proj_kblocks = kbp_ptr.back();
    h->info.proj_kblocks_dense = (int64_t)(TP / GEMM_BM) * (KCP / GEMM_BK);

    // ctc (KS/template_matching.py:110-121) re-laid as ctc_i[s'][t'][l] (LP-padded) in internal
    // order, rows padded with zeros to the register tile height of the peel kernel
    h->ctc_rows = T;
    for (int R : {8, 16, 32, 64})
        if (T <= REG_MAX_T && (T + REG_NRG - 1) / REG_NRG <= R) { h->ctc_rows = REG_NRG * R; break; }
    h->ctc_rows = std::max(h->ctc_rows, h->min_rows);
    const int TS = h->ctc_rows;
    // Without ctc.npy the table is computed on the GPU once the templates are resident
    // (launch_ctc, FP64 accumulation); here only WtW and the channel spans are prepared.
    const int L = 2 * M + 1;
    const bool ctc_file = file_exists(dir + "ctc.npy");
    std::vector<float> ctc_i;
    std::vector<int> tlo(T, 0), thi(T, 0);
    std::vector<double> wtw;
    if (ctc_file) {
        ctc_i.assign((size_t)T * TS * LP, 0.f);
        std::vector<float> ctc;
        auto a = npy_load(dir + "ctc.npy");
        expect_shape(a, {T_file, T_file, L}, "ctc.npy");
        if (a.kind == 'f' && a.itemsize == 4) {
            ctc.resize((size_t)a.numel());
            std::memcpy(ctc.data(), a.bytes.data(), a.bytes.size());
        } else {
            ctc = a.as<float>();
        }
        for (int sp = 0; sp < T; ++sp) {
            const int s = keep.empty() ? h->orig[sp] : keep[h->orig[sp]];
            int lo = T, hi = 0;
            for (int tp = 0; tp < T; ++tp) {
                const int t = keep.empty() ? h->orig[tp] : keep[h->orig[tp]];
                const float* src = &ctc[((size_t)t * T_file + s) * L];
                float* dst = &ctc_i[((size_t)sp * TS + tp) * LP];
                // rows that matter: nonzero, or (sparse level 2) above 2^-40 sqrt(nm_s nm_t)
                const double cut = tau * std::sqrt(nmh[sp] * nmh[tp]);
                bool nz = false;
                for (int l = 0; l < L; ++l) {
                    dst[l] = src[l];
                    nz |= std::fabs((double)src[l]) > cut;
                }
                if (nz) { lo = std::min(lo, tp); hi = tp + 1; }
            }
            tlo[sp] = lo;
            thi[sp] = hi;
        }
    } else {
        // WtW[k1][k2][l] = sum_m W[k1][m - (l - M)] W[k2][m] (KS/template_matching.py:113-114)
        wtw.assign((size_t)K * K * L, 0.0);
        for (int k1 = 0; k1 < K; ++k1)
            for (int k2 = 0; k2 < K; ++k2)
                for (int l = 0; l < L; ++l) {
                    const int d = l - M;
                    double s = 0;
                    for (int m = 0; m < M; ++m) {
                        const int mi = m - d;
                        if (mi >= 0 && mi < M) s += wpca[(size_t)k1 * M + mi] * wpca[(size_t)k2 * M + m];
                    }
                    wtw[((size_t)k1 * K + k2) * L + l] = s;
                }
    }

    // G[k'][k][d + M - 1] = sum_m W[k'][m + d] W[k][m]: PC projection of a shifted waveform,
    // used to take the residual's subtracted spikes out of the features.
    std::vector<float> Gt((size_t)K * K * (2 * M - 1), 0.f);
    for (int k1 = 0; k1 < K; ++k1)
        for (int k2 = 0; k2 < K; ++k2)
            for (int d = -(M - 1); d <= M - 1; ++d) {
                double s = 0;
                for (int m = 0; m < M; ++m) {
                    const int mi = m + d;
                    if (mi >= 0 && mi < M) s += wpca[(size_t)k1 * M + mi] * wpca[(size_t)k2 * M + m];
                }
                Gt[((size_t)k1 * K + k2) * (2 * M - 1) + d + M - 1] = (float)s;
            }

    std::vector<int> iCCi((size_t)nC * C), iUi(T);
    for (int j = 0; j < nC; ++j)
        for (int c = 0; c < C; ++c) iCCi[(size_t)j * C + c] = (int)iCC[(size_t)j * C + c];
    for (int tp = 0; tp < T; ++tp) iUi[tp] = (int)iU[h->orig[tp]];
    std::vector<float> wpf((size_t)K * M);
    for (size_t i = 0; i < wpf.size(); ++i) wpf[i] = (float)wpca[i];

    // ---- device allocation ----
    HIPCHK(hipSetDevice(h->dev));
    const int Wmax = h->Wmax;
    h->BS = (int)round_up(round_up(Wmax, GEMM_BN), 256) + 256;
    h->XS = h->BS + 2 * XOFF + 64;
    h->d_Wr = dalloc<float>(Wrp.size()); upload(h->d_Wr, Wrp);
    h->d_U2 = dalloc<float>(U2.size()); upload(h->d_U2, U2);
    if (h->bf16) {
        std::vector<uint16_t> U2b(U2.size());
        for (size_t i = 0; i < U2.size(); ++i) U2b[i] = bf16_bits(U2[i]);
        h->d_U2b = dalloc<uint16_t>(U2b.size()); upload(h->d_U2b, U2b);
    }
    if (h->x3) {
        // u = hi + lo + e: hi = bf16(u), lo = bf16(u - hi), both rounded to nearest even
        auto split = [](const std::vector<float>& u, std::vector<uint16_t>& hi, std::vector<uint16_t>& lo) {
            hi.resize(u.size());
            lo.resize(u.size());
            for (size_t i = 0; i < u.size(); ++i) {
                hi[i] = bf16_bits(u[i]);
                lo[i] = bf16_bits(u[i] - bf16_value(hi[i]));
            }
        };
        std::vector<uint16_t> hi, lo;
        split(U2, hi, lo);
        h->d_U2b = dalloc<uint16_t>(hi.size()); upload(h->d_U2b, hi);
        h->d_U2l = dalloc<uint16_t>(lo.size()); upload(h->d_U2l, lo);
        split(Wrp, hi, lo);
        h->d_Wrh = dalloc<uint16_t>(hi.size()); upload(h->d_Wrh, hi);
        h->d_Wrl = dalloc<uint16_t>(lo.size()); upload(h->d_Wrl, lo);
        // the PC conv's B operand Q[n][j'] = wpca[n / 16][j' - n % 16 + M/2 - 32] (zero outside
        // 0 .. M - 1 and for n / 16 >= K), n < 32 ceil(K / 2), j' < PCX_K: hi rows, then lo rows
        const int nq = 32 * ((K + 1) / 2);
        std::vector<float> q((size_t)nq * PCX_K, 0.f);
        for (int n = 0; n < nq; ++n)
            for (int jp = 0; jp < PCX_K; ++jp) {
                const int kq = n / 16, m = jp - n % 16 + M / 2 - 32;
                if (kq < K && m >= 0 && m < M) q[(size_t)n * PCX_K + jp] = wpf[(size_t)kq * M + m];
            }
        split(q, hi, lo);
        hi.insert(hi.end(), lo.begin(), lo.end());
        h->d_pcq = dalloc<uint16_t>(hi.size()); upload(h->d_pcq, hi);
    }
    h->d_U = dalloc<float>(Ui.size()); upload(h->d_U, Ui);
    h->d_wpca = dalloc<float>(wpf.size()); upload(h->d_wpca, wpf);
    h->d_ctc = dalloc<float>((size_t)T * TS * LP);
    h->d_tlo = dalloc<int>(T);
    h->d_thi = dalloc<int>(T);
    if (ctc_file) {
        upload(h->d_ctc, ctc_i);
    } else {
        // nonzero channel span of each template (internal order)
        std::vector<int> span(2 * (size_t)T);
        for (int tp = 0; tp < T; ++tp) {
            int c0 = C, c1 = 0;
            for (int k = 0; k < K; ++k)
                for (int c = 0; c < C; ++c)
                    if (Ui[((size_t)tp * K + k) * C + c] != 0.f) { c0 = std::min(c0, c); c1 = std::max(c1, c + 1); }
            span[2 * tp] = c0;
            span[2 * tp + 1] = std::max(c0, c1);
        }
        double* d_wtw = dalloc<double>(wtw.size());
        int* d_span = dalloc<int>(span.size());
        double* d_nmh = dalloc<double>(nmh.size());
        upload(d_wtw, wtw);
        upload(d_span, span);
        upload(d_nmh, nmh);
        HIPCHK(hipMemset(h->d_ctc, 0, (size_t)T * TS * LP * sizeof(float)));
        launch_ctc(h->d_U, d_wtw, reinterpret_cast<const int2*>(d_span), T, K, C, M, TS, h->d_ctc, h->d_tlo,
                   h->d_thi, d_nmh, tau, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(tlo.data(), h->d_tlo, T * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(thi.data(), h->d_thi, T * sizeof(int), hipMemcpyDeviceToHost));
        (void)hipFree(d_wtw);
        (void)hipFree(d_span);
        (void)hipFree(d_nmh);
    }
    for (int sp = 0; sp < T; ++sp) {
        if (!h->cfg.sparse) { tlo[sp] = 0; thi[sp] = T; }
        if (tlo[sp] > thi[sp]) tlo[sp] = thi[sp] = 0;
    }
    upload(h->d_tlo, tlo);
    upload(h->d_thi, thi);
    h->d_Gt = dalloc<float>(Gt.size()); upload(h->d_Gt, Gt);
    {
        std::vector<int> orig_p(TS, INT32_MAX);
        std::copy(h->orig.begin(), h->orig.end(), orig_p.begin());
        h->d_orig = dalloc<int>(TS); upload(h->d_orig, orig_p);
        for (int tp = 0; tp < T; ++tp)
            if (h->orig[tp] == 0) h->t_first = tp;
    }
    h->d_iCC = dalloc<int>(iCCi.size()); upload(h->d_iCC, iCCi);
    h->d_iU = dalloc<int>(T); upload(h->d_iU, iUi);
    h->iU_int = iUi;
    h->d_cmap = dalloc<int>(CP);
    h->cmap_cur.assign(CP, -1);            // nothing uploaded yet
    h->d_kbw_ptr = dalloc<int>(kbw_ptr.size()); upload(h->d_kbw_ptr, kbw_ptr);
    h->d_kbw_idx = dalloc<int>(kbw_idx.size()); upload(h->d_kbw_idx, kbw_idx);
    h->d_kbp_ptr = dalloc<int>(kbp_ptr.size()); upload(h->d_kbp_ptr, kbp_ptr);
    h->d_kbp_idx = dalloc<int>(kbp_idx.size()); upload(h->d_kbp_idx, kbp_idx);
    // norms padded past T: +inf (Cf = 0) and reciprocal 0, filled in below T by k_norms
    h->d_nm = dalloc<float>(TS); upload(h->d_nm, std::vector<float>(TS, INFINITY));
    h->d_rnm = dalloc<float>(TS);
    h->d_nm64 = dalloc<double>(T);
    h->d_Xw = dalloc<float>((size_t)CP * h->XS);
    h->d_Bpc = dalloc<float>((size_t)KCP * h->BS);
    h->d_B = dalloc<float>((size_t)TP * h->BS);
    if (h->bf16) {
        h->d_Bpcb = dalloc<uint16_t>((size_t)KCP * h->BS);
        h->d_Binit = dalloc<float>((size_t)TP * h->BS);
        h->d_Nsq = dalloc<float>((size_t)(TP / GEMM_BM) * h->BS);
        h->d_act = dalloc<uint8_t>((size_t)(h->BS / GEMM_BN) * (TP / GEMM_BM));
        h->d_rescnt = dalloc<unsigned long long>(2);
        HIPCHK(hipMemset(h->d_rescnt, 0, 2 * sizeof(unsigned long long)));
    }
    h->d_Cfmax = dalloc<float>(Wmax);
    {
        const int Rs[5] = {0, 8, 16, 32, 64};
        for (int i = 0; i < 5; ++i) {
            h->peel_cap[i] = peel_max_blocks(h->dev, Rs[i]);
            if (h->peel_cap[i] <= 0) throw LssError(LSS_ERR_HIP, "cannot query peel kernel occupancy");
        }
        // A profiler that cannot follow cooperative dispatches can ask for a plain
        // launch; the grid is still sized to be co-resident on an idle device.
        const char* lm = std::getenv("LSS_PEEL_LAUNCH");
        h->coop = !(lm && std::string(lm) == "plain");
    }
    h->max_groups = (Wmax + MIN_CW - 1) / MIN_CW;
    h->slot_cap = 256;
    h->cap_g = MAX_LIST;
    h->d_slots = dalloc<Spike>((size_t)h->max_groups * h->slot_cap);
    h->d_list = dalloc<Spike>((size_t)h->max_groups * h->cap_g);
    h->d_slot_cnt = dalloc<int>(h->max_groups);
    // split the templates of a column block over up to 16 workgroups, 128 rows each at most
    if (h->T <= REG_MAX_T && h->sem == LSS_SEM_KILOSORT && h->min_rows == 0) {
        h->max_split = std::min<int>(16, (h->T + 127) / 128);
        if (h->cfg.peel_split > 0) h->max_split = std::min(h->max_split, (int)h->cfg.peel_split);
    } else if (h->min_rows > 0 && h->group_split > 1) {
        h->max_split = h->group_split;   // the group chose this stream's split
    }
    // per-subset column maxima (score, B value, template), one copy per iteration
    // parity, as tagged words; all zero (a tag no epoch uses) to start with
    h->ptag_words = 3 * 2 * (int64_t)std::max(1, h->max_split) * Wmax;
    h->d_ptag = dalloc<unsigned long long>(h->ptag_words);
    HIPCHK(hipMemset(h->d_ptag, 0, h->ptag_words * sizeof(unsigned long long)));
    h->d_epoch = dalloc<unsigned>(1);
    if (const char* es = std::getenv("LSS_EPOCH_START")) {
        // test support: start the tag epoch elsewhere (near its wrap-around)
        const unsigned e0 = (unsigned)std::strtoul(es, nullptr, 10) & EPOCH_MASK;
        HIPCHK(hipMemcpy(h->d_epoch, &e0, sizeof(unsigned), hipMemcpyHostToDevice));
    }
    h->d_list_cnt = dalloc<int>(h->max_groups);
    // per-iteration spike totals and done counters, then the peel's neighbour flags
    // (per workgroup, per column block); all cleared by the whitening launch each batch
    h->g_cap = h->max_groups * std::max(1, h->max_split);
    h->sync_words = 2 * std::max(1, h->max_iter) + h->g_cap + h->max_groups;
    h->d_iter_cnt = dalloc<int>(h->sync_words);
    h->d_err = dalloc<int>(1);
    h->d_bar = dalloc<unsigned>(BAR_WORDS);
    h->d_prof = dalloc<unsigned long long>(PROF_WORDS);
    if (h->cfg.keep_stages) {
        h->d_dbgB = dalloc<float>((size_t)TP * h->BS);
        h->d_dbgCf = dalloc<float>(Wmax);
        h->d_dbgIm = dalloc<int>(Wmax);
    }
    // preprocessing (cfg.preprocess): hp_filter.npy (loadKilosortTrainingData, OnlineSpikesV2.cu:597-616)
    h->pre = h->cfg.preprocess;
    if (h->pre & ~(PRE_MEAN | PRE_CAR | PRE_HP)) throw LssError(LSS_ERR_ARG, "unknown preprocessing bits");
    if (h->pre) {
        if (C > PRE_MAX_C) throw LssError(LSS_ERR_ARG, "preprocessing supports at most 1024 channels");
        if (h->pre & PRE_HP) {
            if (!file_exists(dir + "hp_filter.npy"))
                throw LssError(LSS_ERR_IO, "high-pass filtering needs hp_filter.npy");
            auto hp = npy_load(dir + "hp_filter.npy").as<float>();
            if (hp.empty()) throw LssError(LSS_ERR_IO, "hp_filter.npy is empty");
            h->hp_F = (int)hp.size();
            // the span of taps that matter: the response's tails outside [hp_lo, hp_hi) hold
            // at most 2^-30 of its summed magnitude (each side 2^-31)
            double tot = 0;
            for (float v : hp) tot += std::fabs((double)v);
            const double eps = tot * std::ldexp(1.0, -31);
            double acc = 0;
            int lo = 0;
            while (lo < h->hp_F && acc + std::fabs((double)hp[lo]) <= eps) acc += std::fabs((double)hp[lo++]);
            acc = 0;
            int hi = h->hp_F;
            while (hi > lo && acc + std::fabs((double)hp[hi - 1]) <= eps) acc += std::fabs((double)hp[--hi]);
            h->hp_lo = lo;
            h->hp_hi = hi;
            h->d_hp = dalloc<float>(hp.size()); upload(h->d_hp, hp);
        }
        h->PS = round_up(Wmax, 64);
        const int nslab = (Wmax + PRE_SLAB - 1) / PRE_SLAB;
        h->d_psum = dalloc<int>((size_t)nslab * C);
        h->d_pmin = dalloc<int>((size_t)nslab * C);
        h->d_pmax = dalloc<int>((size_t)nslab * C);
        h->d_pmean = dalloc<float>(C);
        h->d_pmed = dalloc<float>(Wmax);
        h->d_X1 = dalloc<float>((size_t)C * h->PS);
        if (h->pre & PRE_HP) h->d_Xp = dalloc<float>((size_t)C * h->PS);
    }
    // probe geometry for spike positions, final-cluster centroids for relabelling
    if (file_exists(dir + "xc.npy") && file_exists(dir + "yc.npy")) {
        auto xa = npy_load(dir + "xc.npy"), ya = npy_load(dir + "yc.npy");
        if (xa.numel() != C || ya.numel() != C) throw LssError(LSS_ERR_IO, "xc/yc must have C entries");
        if (nC > 16) throw LssError(LSS_ERR_IO, "spike positions support at most 16 nearest channels");
        h->d_xc = dalloc<float>(C); upload(h->d_xc, xa.as<float>());
        h->d_yc = dalloc<float>(C); upload(h->d_yc, ya.as<float>());
    }
    if (h->cfg.relabel) {
        if (!h->d_xc) throw LssError(LSS_ERR_IO, "relabelling needs xc.npy and yc.npy");
        // cluster centres as (x, y) pairs (loadKilosortClusteringData, OnlineSpikesV2.cu:648-659)
        auto ca = npy_load(dir + "cluster_centroids.npy");
        if (ca.numel() < 2 || ca.numel() % 2) throw LssError(LSS_ERR_IO, "cluster_centroids.npy must hold (x, y) pairs");
        const auto cv = ca.as<float>();
        h->n_cent = (int)(cv.size() / 2);
        h->d_cent = dalloc<float>(cv.size()); upload(h->d_cent, cv);
    }
    // mapped, coherent host memory: the peel kernel writes events straight into it
    h->ev_cap = h->cfg.max_events;
    const size_t out_bytes = sizeof(OutHeader) + sizeof(lss_event) * (size_t)h->ev_cap +
                             sizeof(float) * (size_t)h->ev_cap * nC * K + sizeof(float) * 2 * (size_t)h->ev_cap;
    HIPCHK(hipHostMalloc(&h->h_out, out_bytes, hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(h->h_out, 0, out_bytes);
    void* dout = nullptr;
    HIPCHK(hipHostGetDevicePointer(&dout, h->h_out, 0));
    h->hdr = static_cast<OutHeader*>(h->h_out);
    h->ev = reinterpret_cast<lss_event*>(h->hdr + 1);
    h->feat = reinterpret_cast<float*>(h->ev + h->ev_cap);
    h->d_hdr = static_cast<OutHeader*>(dout);
    h->d_ev = reinterpret_cast<lss_event*>(h->d_hdr + 1);
    h->d_feat = reinterpret_cast<float*>(h->d_ev + h->ev_cap);
    if (h->d_xc) {
        h->pos = h->feat + (size_t)h->ev_cap * nC * K;
        h->d_pos = h->d_feat + (size_t)h->ev_cap * nC * K;
    }

    HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    if (h->cfg.profile)
        for (auto& e : h->evt) HIPCHK(hipEventCreate(&e));
    launch_norms(h->d_U, K * C, T, h->d_nm, h->d_rnm, h->d_nm64, h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    h->last_spike.assign(std::max(T, h->n_cent), INT64_MIN / 2);   // indexed by template, or by cluster (relabel)
}

// Validate the channel map of a call and upload it when it changed.  cmap_cur
// holds CP entries from load time on (-1: nothing uploaded yet), so a call with
// the same map touches no heap memory.
void set_cmap(lss_handle* h, const int32_t* chan_map, int32_t ld) {
    bool same = true;
    for (int c = 0; c < h->C; ++c) {
        const int v = chan_map ? chan_map[c] : h->cmap_default[c];
        if (v < 0 || v >= ld) throw LssError(LSS_ERR_ARG, "channel map entry out of range of the row stride");
        same = same && h->cmap_cur[c] == v;
    }
    if (same) return;
    for (int c = 0; c < h->C; ++c) h->cmap_cur[c] = chan_map ? chan_map[c] : h->cmap_default[c];
    for (int c = h->C; c < h->CP; ++c) h->cmap_cur[c] = 0;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(h->d_cmap, h->cmap_cur.data(), (size_t)h->CP * sizeof(int), hipMemcpyHostToDevice));
    h->cmap_id = true;
    for (int c = 0; c < h->C; ++c) h->cmap_id = h->cmap_id && h->cmap_cur[c] == c;
}

// Run one framed batch that is already on the device.
}  // namespace

// The per-batch device buffers a batch runs on: the handle's own (allocated
// with it), or a region of the static scratch array (the allocation-free entry
// point).  The tables (templates, ctc, norms, ...) always come from the handle.
struct BatchBufs {
    float *Xw = nullptr, *Bpc = nullptr, *B = nullptr, *Binit = nullptr, *Nsq = nullptr;
    uint8_t* act = nullptr;
    float* Cfmax = nullptr;
    uint16_t* Bpcb = nullptr;
    Spike *slots = nullptr, *list = nullptr;
    int *slot_cnt = nullptr, *list_cnt = nullptr, *iter_cnt = nullptr, *err = nullptr;
    unsigned* bar = nullptr;
    unsigned long long *prof = nullptr, *ptag = nullptr;
    int64_t ptag_words = 0;
    unsigned* epoch = nullptr;
    OutHeader* hdr = nullptr;          // device address of the output header, events, features
    lss_event* ev = nullptr;
    float *feat = nullptr, *pos = nullptr;
    int *psum = nullptr, *pmin = nullptr, *pmax = nullptr;
    float *pmean = nullptr, *pmed = nullptr, *X1 = nullptr, *Xp = nullptr;
    float *dbgB = nullptr, *dbgCf = nullptr;
    int* dbgIm = nullptr;
    uint8_t* gate = nullptr;           // proposal gate: the score tiles to compute
};

BatchBufs handle_bufs(const lss_handle* h) {
    BatchBufs b;
    b.Xw = h->d_Xw; b.Bpc = h->d_Bpc; b.B = h->d_B; b.Binit = h->d_Binit; b.Nsq = h->d_Nsq; b.act = h->d_act;
    b.Cfmax = h->d_Cfmax;
    b.Bpcb = h->d_Bpcb;
    b.slots = h->d_slots; b.list = h->d_list;
    b.slot_cnt = h->d_slot_cnt; b.list_cnt = h->d_list_cnt; b.iter_cnt = h->d_iter_cnt; b.err = h->d_err;
    b.bar = h->d_bar; b.prof = h->d_prof; b.ptag = h->d_ptag; b.ptag_words = h->ptag_words; b.epoch = h->d_epoch;
    b.hdr = h->d_hdr; b.ev = h->d_ev; b.feat = h->d_feat; b.pos = h->d_pos;
    b.psum = h->d_psum; b.pmin = h->d_pmin; b.pmax = h->d_pmax;
    b.pmean = h->d_pmean; b.pmed = h->d_pmed; b.X1 = h->d_X1; b.Xp = h->d_Xp;
    if (h->cfg.keep_stages) { b.dbgB = h->d_dbgB; b.dbgCf = h->d_dbgCf; b.dbgIm = h->d_dbgIm; }
    b.gate = h->d_gate;
    return b;
}

// The handle's per-batch device buffers laid out from base: the tag epoch at base,
// the tagged column maxima after it (tag_bytes in all), then the rest; *need = bytes
// used.  base may be null to size the layout.
BatchBufs layout_bufs(const lss_handle* h, unsigned char* base, size_t tag_bytes, size_t* need) {
    BatchBufs b;
    b.epoch = reinterpret_cast<unsigned*>(base);
    b.ptag = reinterpret_cast<unsigned long long*>(base + 256);
    b.ptag_words = (int64_t)((tag_bytes - 256) / sizeof(unsigned long long));
    size_t off = tag_bytes;
    auto take = [&](size_t bytes) {
        off = (off + 255) / 256 * 256;
        unsigned char* p = base + off;
        off += bytes;
        return p;
    };
    const size_t CP = h->CP, TP = h->TP, KCP = h->KCP, BS = h->BS, XS = h->XS, Wmax = h->Wmax;
    b.Xw = reinterpret_cast<float*>(take(CP * XS * 4));
    b.Bpc = reinterpret_cast<float*>(take(KCP * BS * 4));
    b.B = reinterpret_cast<float*>(take(TP * BS * 4));
    if (h->bf16) {
        b.Bpcb = reinterpret_cast<uint16_t*>(take(KCP * BS * 2));
        b.Binit = reinterpret_cast<float*>(take(TP * BS * 4));
        b.Nsq = reinterpret_cast<float*>(take(TP / GEMM_BM * BS * 4));
        b.act = reinterpret_cast<uint8_t*>(take(BS / GEMM_BN * (TP / GEMM_BM)));
    }
    b.Cfmax = reinterpret_cast<float*>(take(Wmax * 4));
    b.slots = reinterpret_cast<Spike*>(take((size_t)h->max_groups * h->slot_cap * sizeof(Spike)));
    b.list = reinterpret_cast<Spike*>(take((size_t)h->max_groups * h->cap_g * sizeof(Spike)));
    b.slot_cnt = reinterpret_cast<int*>(take((size_t)h->max_groups * 4));
    b.list_cnt = reinterpret_cast<int*>(take((size_t)h->max_groups * 4));
    b.iter_cnt = reinterpret_cast<int*>(take((size_t)h->sync_words * 4));
    b.err = reinterpret_cast<int*>(take(4));
    b.bar = reinterpret_cast<unsigned*>(take((size_t)BAR_WORDS * 4));
    b.prof = reinterpret_cast<unsigned long long*>(take((size_t)PROF_WORDS * 8));
    const size_t fw = (size_t)h->nC * h->K;
    b.hdr = reinterpret_cast<OutHeader*>(take(sizeof(OutHeader)));
    b.ev = reinterpret_cast<lss_event*>(take((size_t)h->ev_cap * sizeof(lss_event)));
    b.feat = reinterpret_cast<float*>(take((size_t)h->ev_cap * fw * 4));
    if (h->d_xc) b.pos = reinterpret_cast<float*>(take((size_t)h->ev_cap * 2 * 4));
    if (h->pre) {
        const size_t nslab = (Wmax + PRE_SLAB - 1) / PRE_SLAB, C = h->C, PS = h->PS;
        b.psum = reinterpret_cast<int*>(take(nslab * C * 4));
        b.pmin = reinterpret_cast<int*>(take(nslab * C * 4));
        b.pmax = reinterpret_cast<int*>(take(nslab * C * 4));
        b.pmean = reinterpret_cast<float*>(take(C * 4));
        b.pmed = reinterpret_cast<float*>(take(Wmax * 4));
        b.X1 = reinterpret_cast<float*>(take(C * PS * 4));
        if (h->pre & PRE_HP) b.Xp = reinterpret_cast<float*>(take(C * PS * 4));
    }
    *need = off;
    return b;
}

struct BatchPlan {
    PreJob pre;
    bool has_pre = false;
    WhitenJob wj;
    PcconvJob cj;
    ProjectJob pj;
    PeelParams pp;
    int R = 0, CW = 0, G = 0;
};

// PC conv workgroups per 256-column block: channel rows PC_ROWS_PER_BLOCK at a time.

inline int peel_cap_index(int R) { return R == 0 ? 0 : R == 8 ? 1 : R == 16 ? 2 : R == 32 ? 3 : 4; }

// Job descriptors of one handle for one batch of W columns.  The input is read
// from d_x, or, when xptr is set, from the device pointer stored at *xptr.
void plan_batch(lss_handle* h, const BatchBufs& bb, const int16_t* d_x, const int16_t* const* xptr, int W, int64_t ld,
                int CW, int R, int S, BatchPlan& bp) {
    const int M = h->M;
    WhitenJob& wj = bp.wj;
    wj = WhitenJob{};
    wj.Wr = h->d_Wr; wj.x = d_x; wj.xptr = xptr; wj.cmap = h->d_cmap;
    wj.kb_ptr = h->d_kbw_ptr; wj.kb_idx = h->d_kbw_idx; wj.Xw = bb.Xw;
    wj.ld = ld; wj.CP = h->CP; wj.W = W; wj.C = h->C; wj.XS = h->XS;
    wj.mtiles = h->CP / GEMM_BM;
    wj.cmap_id = h->cmap_id ? 1 : 0;
    if (h->x3) {
        wj.Wrh = reinterpret_cast<const __bf16*>(h->d_Wrh);
        wj.Wrl = reinterpret_cast<const __bf16*>(h->d_Wrl);
    }
    wj.ntiles = (W + XOFF + GEMM_BN - 1) / GEMM_BN;          // cover the right zero pad
    bp.has_pre = h->pre != 0;
    if (bp.has_pre) {
        PreJob& q = bp.pre;
        q = PreJob{};
        q.x = d_x; q.xptr = xptr; q.ld = ld; q.cmap = h->d_cmap;
        q.C = h->C; q.W = W; q.flags = h->pre; q.nslab = (W + PRE_SLAB - 1) / PRE_SLAB;
        q.psum = bb.psum; q.pmin = bb.pmin; q.pmax = bb.pmax;
        q.mean = bb.pmean; q.med = bb.pmed; q.X1 = bb.X1; q.Xp = bb.Xp; q.PS = h->PS;
        q.p2p = &bb.hdr->p2p;
        if (h->pre & PRE_HP) {
            // taps h[m] = hp[m + crop], m in [0, W): the response cropped (F >= W) or zero-padded
            // to W (fft_highpass, KS/preprocessing.py:143-153); only m with hp index in
            // [hp_lo, hp_hi) are kept.  Kilosort's fftshift reads y[(k - W/2) mod W], the
            // reference's shift_kernel y[(k + W/2) mod W] (myGPUhelpers.cu:72-83).
            const int F = h->hp_F;
            if (h->sem == LSS_SEM_LSS && W > F)
                throw LssError(LSS_ERR_ARG, "batch of " + std::to_string(W) + " samples is longer than the " +
                                            std::to_string(F) + "-sample high-pass filter");
            const int crop = F >= W ? (F - W) / 2 : -((W - F) / 2);
            const int m_lo = std::max(0, h->hp_lo - crop), m_hi = std::min(W, h->hp_hi - crop);
            const int s0 = h->sem == LSS_SEM_KILOSORT ? (W - W / 2) % W : (W / 2) % W;
            q.ntap = std::max(0, m_hi - m_lo);
            q.taps = h->d_hp + crop + m_lo;
            q.shift = (int)(((int64_t)s0 + std::max(0, m_lo)) % W);
        }
        wj.xf = (h->pre & PRE_HP) ? bb.Xp : bb.X1;
        wj.xfs = h->PS;
    }
    PcconvJob& cj = bp.cj;
    cj = PcconvJob{};
    cj.Xw = bb.Xw; cj.wpca = h->d_wpca; cj.Bpc = bb.Bpc;
    cj.Bpcb = reinterpret_cast<__bf16*>(bb.Bpcb);
    cj.active = xptr;
    cj.XS = h->XS; cj.M = M; cj.BS = h->BS; cj.CP = h->CP; cj.C = h->C; cj.x3 = h->x3 ? 1 : 0;
    cj.q = reinterpret_cast<const __bf16*>(h->d_pcq);
    cj.ncols = (W + GEMM_BN - 1) / GEMM_BN * GEMM_BN;
    cj.wblocks = ((W + GEMM_BN - 1) / GEMM_BN * GEMM_BN + 255) / 256;
    ProjectJob& pj = bp.pj;
    pj = ProjectJob{};
    pj.A = h->d_U2; pj.Bm = bb.Bpc; pj.kb_ptr = h->d_kbp_ptr; pj.kb_idx = h->d_kbp_idx; pj.C = bb.B;
    pj.lda = h->KCP; pj.ldb = h->BS; pj.ldc = h->BS;
    if (h->prop) { pj.gate = bb.gate; pj.gate_mt = h->TP / GEMM_BM; }
    pj.active = xptr;
    if (h->x3) {
        pj.Ab = reinterpret_cast<const __bf16*>(h->d_U2b);
        pj.Al = reinterpret_cast<const __bf16*>(h->d_U2l);
    }
    if (h->bf16) {
        pj.Ab = reinterpret_cast<const __bf16*>(h->d_U2b);
        pj.Bmb = reinterpret_cast<const __bf16*>(bb.Bpcb);
        pj.C2 = bb.Binit;
        pj.Nsq = bb.Nsq;
        // tiles that could reach lo2 are recomputed by the FP32 projection
        pj.act = bb.act;
        pj.rnm = h->d_rnm;
        pj.Th2 = h->Th * h->Th;
        pj.lo2 = 0.9f * pj.Th2;             // as the peel's
        pj.n_rescored = h->d_rescnt;
    }
    pj.mtiles = h->TP / GEMM_BM;
    pj.ntiles = (W + GEMM_BN - 1) / GEMM_BN;
    PeelParams& p = bp.pp;
    p = PeelParams{};
    const int NB = (W + CW - 1) / CW;
    const int G = S * NB;
    if (NB > h->max_groups || G > h->g_cap) throw LssError(LSS_ERR_ARG, "peel grid larger than the handle's flag space");
    p.B = bb.B; p.BS = h->BS;
    p.B_bytes = (int)std::min<int64_t>((int64_t)h->TP * h->BS * 4, INT32_MAX);
    p.ctc_bytes = (int)std::min<int64_t>((int64_t)h->T * h->ctc_rows * LP * 4, INT32_MAX);
    p.ctc_rows = h->ctc_rows;
    p.t_first = h->t_first;
    p.T = h->T; p.W = W; p.M = M;
    p.nm = h->d_nm; p.rnm = h->d_rnm;
    p.ctc = h->d_ctc; p.tlo = h->d_tlo; p.thi = h->d_thi; p.orig = h->d_orig;
    p.Th2 = h->Th * h->Th;
    p.max_iter = h->max_iter;
    p.CW = CW; p.G = G;
    p.S = S; p.NB = NB;
    p.Pa = bb.ptag; p.epoch = bb.epoch;
    p.lss = h->sem == LSS_SEM_LSS;
    p.Cfmax = bb.Cfmax;
    p.slots = bb.slots; p.slot_cnt = bb.slot_cnt; p.slot_cap = h->slot_cap;
    p.list = bb.list; p.list_cnt = bb.list_cnt; p.cap_g = h->cap_g;
    p.bar = bb.bar; p.iter_cnt = bb.iter_cnt; p.err = bb.err;
    p.Bpc = bb.Bpc; p.CP = h->CP;
    p.U = h->d_U; p.C = h->C; p.K = h->K; p.nC = h->nC;
    p.iCC = h->d_iCC; p.iU = h->d_iU; p.Gt = h->d_Gt;
    p.do_feat = 1;
    p.hdr = bb.hdr; p.ev = bb.ev; p.feat = bb.feat; p.ev_cap = h->ev_cap;
    p.pos = bb.pos; p.xc = h->d_xc; p.yc = h->d_yc;
    p.cent = h->d_cent; p.n_cent = h->n_cent;
    p.half = M / 2; p.nt0min = h->nt0min;
    p.dbg_cfmax = bb.dbgCf;
    p.dbg_imax = bb.dbgIm;
    p.prof = bb.prof;
    p.clock = h->cfg.profile == 1;     // the phase clock: profile level 1 only
    if (h->prop) { p.gate = bb.gate; p.gate_mt = h->TP / GEMM_BM; }
    p.active = xptr;
    if (h->bf16) {
        // candidates within 10% of threshold on the score scale are decided in FP32
        p.rescore = 1;
        p.lo2 = 0.9f * p.Th2;
        p.U2 = h->d_U2; p.ldu = h->KCP;
        p.kb_ptr = h->d_kbp_ptr; p.kb_idx = h->d_kbp_idx;
        p.Binit = bb.Binit;
        p.Nsq = bb.Nsq;
        p.act = bb.act; p.act_mt = h->TP / GEMM_BM;
        p.n_rescored = h->d_rescnt;
    }
    bp.R = R; bp.CW = CW; bp.G = G;
    (void)R;
}

// Peel geometry for one handle.  Register tiles (64 columns) first, with the
// templates split over S workgroups per column block when the co-resident grid
// has room: 16 rows per thread when that fits, else the fewest rows that fit.
// Otherwise the narrowest column block whose grid is co-resident.
// Register rows per thread for S subsets of T templates (0: no register tile fits).
int split_rows(int T, int S) {
    for (int r : {8, 16, 32, 64})
        if (S * REG_NRG * r >= T) return r;
    return 0;
}

// Whether S subsets with R rows each form a valid co-resident split geometry.
bool split_fits(const lss_handle* h, int W, int S, int R) {
    const int NB = (W + MIN_CW - 1) / MIN_CW;
    return S >= 2 && S <= h->max_split && R > 0 && R <= 32 && S * REG_NRG * R >= h->T &&
           S * REG_NRG * R <= h->ctc_rows && NB <= h->max_groups && S * NB <= h->peel_cap[peel_cap_index(R)];
}

// The split sizes worth trying for a batch of W columns: none, then the fewest
// subsets that reach 32, 16 and 8 rows per thread.
int split_candidates(const lss_handle* h, int W, int out[4]) {
    int n = 0;
    out[n++] = 1;
    for (int r : {32, 16, 8}) {
        const int s2 = (h->T + REG_NRG * r - 1) / (REG_NRG * r);
        if (split_fits(h, W, s2, split_rows(h->T, s2)) && std::find(out, out + n, s2) == out + n) out[n++] = s2;
    }
    return n;
}

void peel_geometry(const lss_handle* h, int W, int& CW, int& R, int& S) {
    S = 1;
    int want = h->force_split;
    if (want == 0) {
        const auto it = h->tuned_split.find(W);
        if (it != h->tuned_split.end()) want = it->second;
    }
    if (want == 0 && h->cfg.peel_split > 1) want = h->cfg.peel_split;   // the caller's split, if it fits
    if (want > 1) {
        const int r = split_rows(h->T, want);
        if (split_fits(h, W, want, r)) { CW = MIN_CW; R = r; S = want; return; }
    }
    if (want == 0 && h->max_split > 1) {
        // default: 16 rows per thread when that split fits (measured fastest at 384 ch x
        // 512 templates, 3122 columns), else the fewest rows per thread that fit
        int cand[4];
        const int nc = split_candidates(h, W, cand);
        if (nc > 1) {
            S = cand[nc - 1];
            for (int i = 0; i < nc; ++i)
                if (cand[i] > 1 && split_rows(h->T, cand[i]) == 16) S = cand[i];
            CW = MIN_CW; R = split_rows(h->T, S);
            return;
        }
    }
    for (CW = MIN_CW;; CW *= 2) {
        const int G = (W + CW - 1) / CW;
        R = peel_rows(h->T, CW);
        const int cap = h->peel_cap[peel_cap_index(R)];
        if (G <= std::min(cap, h->max_groups)) return;
        if (CW >= MAX_CW) throw LssError(LSS_ERR_ARG, "batch too large for one co-resident peel launch");
    }
}

// Check the peel kernel's error flags of one handle after a batch.  A wait that
// timed out (workgroups not co-resident) is LSS_ERR_TIMEOUT; a full internal
// spike list is LSS_ERR_LIMIT.  Neither is about the caller's event buffer.
void check_peel(lss_handle* h, const OutHeader& hd, int W) {
    if (hd.err) {
        std::string why;
        if (hd.err & ERR_BARRIER) why += " a cross-workgroup wait timed out (peel grid not co-resident);";
        if (hd.err & ERR_LIST) why += " per-column-block spike list full;";
        if (hd.err & ERR_SLOT) why += " per-iteration spike slots full;";
        if (hd.err & ERR_STAGE) why += " neighbour spike staging full;";
        if (hd.err & ERR_BAND) why += " FP32 rescoring pair list full;";
        throw LssError((hd.err & ERR_BARRIER) ? LSS_ERR_TIMEOUT : LSS_ERR_LIMIT, "peel kernel reported:" + why);
    }
    h->last_W = W;
    h->info.batches += 1;
    h->info.last_iters = hd.n_iter;
    h->info.last_spikes = (int32_t)hd.count;
}

namespace {

// Proposal network on the whitened batch, then the tile mask.
void run_gate(lss_handle* h, int W, hipStream_t s, const int16_t* const* active = nullptr) {
    pn_infer(h->prop->wt, h->pnb, h->d_Xw, h->XS, XOFF, h->C, W, s, active);
    pn_gate(h->pnb, h->gate, W, s, active);
}

// Run one framed batch that is already on the device.
void run_batch(lss_handle* h, const BatchBufs& bb, const int16_t* d_x, int W, int64_t ld, int64_t first_sample,
               bool check = true) {
    if (W <= 0 || W > h->Wmax) throw LssError(LSS_ERR_ARG, "batch of " + std::to_string(W) +
                                              " samples outside 1..max_batch " + std::to_string(h->Wmax));
    hipStream_t s = h->stream;
    const bool prof = h->cfg.profile == 1;         // every stage
    const bool prof_peel = h->cfg.profile != 0;    // the peel (level 2: the peel only)
    int CW = 0, R = 0, S = 1;
    peel_geometry(h, W, CW, R, S);
    BatchPlan bp;
    plan_batch(h, bb, d_x, nullptr, W, ld, CW, R, S, bp);
    h->info.peel_cw = CW; h->info.peel_rows = R; h->info.peel_split = S; h->info.peel_groups = bp.G;
    WhitenBatch wb{};
    wb.n = 1; wb.start[0] = 0; wb.start[1] = gemm_blocks(bp.wj.mtiles, bp.wj.ntiles); wb.job[0] = bp.wj;
    PcconvBatch cb{};
    cb.n = 1; cb.start[0] = 0; cb.start[1] = pc_job_blocks(bp.cj); cb.job[0] = bp.cj;
    ProjectBatch pb{};
    pb.n = 1; pb.start[0] = 0; pb.start[1] = gemm_blocks(bp.pj.mtiles, bp.pj.ntiles); pb.job[0] = bp.pj;
    PeelBatch kb{};
    kb.n = 1; kb.g_start[0] = 0; kb.g_start[1] = bp.G; kb.p[0] = bp.pp;
    // the peel words of this batch are reset by the whitening launch
    ResetBatch rb{};
    rb.n = 1;
    rb.bar[0] = bb.bar; rb.iter_cnt[0] = bb.iter_cnt; rb.n_iter[0] = h->sync_words; rb.err[0] = bb.err;
    rb.prof[0] = bb.prof; rb.epoch[0] = bb.epoch; rb.ptag[0] = bb.ptag; rb.ptag_words[0] = bb.ptag_words;
    if (prof) HIPCHK(hipEventRecord(h->evt[1], s));
    if (bp.has_pre) {
        PreBatch pre{};
        pre.n = 1;
        pre.job[0] = bp.pre;
        launch_preprocess(pre, s);
    }
    if (prof) HIPCHK(hipEventRecord(h->evt[2], s));
    launch_whiten(wb, &rb, s);
    if (prof) HIPCHK(hipEventRecord(h->evt[3], s));
    if (h->prop) run_gate(h, W, s);
    if (prof) HIPCHK(hipEventRecord(h->evt[4], s));
    launch_pcconv(cb, h->K, s);
    if (prof) HIPCHK(hipEventRecord(h->evt[5], s));
    launch_project(pb, h->cfg.precision, s);
    if (bb.dbgB)
        HIPCHK(hipMemcpyAsync(bb.dbgB, bb.B, sizeof(float) * (size_t)h->TP * h->BS, hipMemcpyDeviceToDevice, s));
    if (prof_peel) HIPCHK(hipEventRecord(h->evt[6], s));
    HIPCHK(hipGetLastError());
    {
        const hipError_t e = launch_peel(kb, R, h->coop, s);
        if (e == hipErrorIllegalState && h->coop)
            // seen when another process on the same GPU holds its own cooperative launches
            throw LssError(LSS_ERR_HIP, std::string("cooperative peel launch refused (") + hipGetErrorString(e) +
                                            "): one process per GPU, or LSS_PEEL_LAUNCH=plain on an otherwise idle GPU");
        HIPCHK(e);
    }
    if (prof_peel) HIPCHK(hipEventRecord(h->evt[7], s));
    HIPCHK(hipStreamSynchronize(s));
    if (prof_peel) {
        float ms = 0.f;
        if (h->copy_timed) {
            HIPCHK(hipEventElapsedTime(&ms, h->evt[0], h->evt[1]));
            h->stage_ms[0] += ms;
        }
        for (int i = prof ? 1 : LSS_NSTAGES - 1; i < LSS_NSTAGES; ++i) {
            HIPCHK(hipEventElapsedTime(&ms, h->evt[i], h->evt[i + 1]));
            h->stage_ms[i] += ms;
        }
        h->timed_batches += 1;
        h->copy_timed = false;
    }
    h->cur_first = first_sample;
    if (check) check_peel(h, *h->hdr, W);
}

// Copy the kernel's events (already sorted by sample) to the caller, applying
// per-template duplicate suppression when enabled.  *n_out = the number of
// events of the batch.  The handle's state (the dedup table, the event count,
// the positions of the last call) changes only when they all fit; otherwise
// the buffer holds the first `cap` of them, the call fails with
// LSS_ERR_CAPACITY and can be repeated with a larger buffer.
void emit_from(lss_handle* h, const OutHeader& hd, const lss_event* src_ev, const float* src_feat,
               const float* src_pos, lss_event* ev, int64_t cap, float* feat, int64_t lo_sample, int64_t hi_sample,
               bool dedup, int64_t* n_out) {
    *n_out = 0;
    if (hd.count > h->ev_cap)
        throw LssError(LSS_ERR_LIMIT, "the batch has " + std::to_string(hd.count) +
                                          " events, more than the handle's max_events " + std::to_string(h->ev_cap));
    const int64_t n = hd.count;
    if (h->pre) h->info.p2p = hd.p2p;
    const int64_t fw = (int64_t)h->nC * h->K;
    // pass 1: which events are reported (dedup against a scratch copy of the last times)
    std::vector<int64_t> pick;
    pick.reserve((size_t)n);
    std::vector<std::pair<int, int64_t>> touched;    // template, its last time after this batch
    {
        std::map<int, int64_t> last;
        for (int64_t i = 0; i < n; ++i) {
            const int64_t s = src_ev[i].sample + h->cur_first;   // the kernel writes samples relative to column 0
            if (s < lo_sample || s >= hi_sample) continue;
            if (dedup) {
                const int t = src_ev[i].template_id;
                auto it = last.find(t);
                const int64_t prev = it != last.end() ? it->second : h->last_spike[t];
                if (s - prev < h->dt) continue;
                last[t] = s;
            }
            pick.push_back(i);
        }
        touched.assign(last.begin(), last.end());
    }
    const int64_t total = (int64_t)pick.size();
    *n_out = total;
    // pass 2: copy out (the first cap of them when they do not all fit); when every event
    // is reported, the features are one block copy
    const int64_t out = std::min(total, cap);
    for (int64_t k = 0; k < out; ++k) {
        const int64_t i = pick[k];
        if (ev) {
            ev[k] = src_ev[i];
            ev[k].sample += h->cur_first;
        }
        if (feat && total != n) std::memcpy(feat + k * fw, src_feat + i * fw, sizeof(float) * fw);
    }
    if (feat && total == n && out > 0) std::memcpy(feat, src_feat, sizeof(float) * fw * out);
    if (total > cap)
        throw LssError(LSS_ERR_CAPACITY, "event buffer holds " + std::to_string(cap) + " of " +
                                             std::to_string(total) + " events");
    for (const auto& tl : touched) h->last_spike[tl.first] = tl.second;
    h->last_pos.clear();
    if (src_pos)
        for (int64_t k = 0; k < total; ++k) {
            h->last_pos.push_back(src_pos[2 * pick[k]]);
            h->last_pos.push_back(src_pos[2 * pick[k] + 1]);
        }
    h->info.events += total;
}

void emit(lss_handle* h, lss_event* ev, int64_t cap, float* feat, int64_t lo_sample, int64_t hi_sample, bool dedup,
          int64_t* n_out) {
    emit_from(h, *h->hdr, h->ev, h->feat, h->pos, ev, cap, feat, lo_sample, hi_sample, dedup, n_out);
}

int16_t* ensure_dx(lss_handle* h, size_t elems) {
    if (elems > h->d_x_elems) {
        if (h->d_x) HIPCHK(hipFree(h->d_x));
        h->d_x = nullptr;
        HIPCHK(hipMalloc(&h->d_x, elems * sizeof(int16_t)));
        h->d_x_elems = elems;
    }
    return h->d_x;
}

void process_host_batch(lss_handle* h, const int16_t* samples, int64_t W, int32_t ld, const int32_t* chan_map,
                        int64_t first_sample) {
    if (!samples || W <= 0) throw LssError(LSS_ERR_ARG, "empty batch");
    if (W > h->Wmax) throw LssError(LSS_ERR_ARG, "batch longer than max_batch");
    set_cmap(h, chan_map, ld);
    int16_t* dx = ensure_dx(h, (size_t)W * ld);
    if (h->cfg.profile == 1) HIPCHK(hipEventRecord(h->evt[0], h->stream));
    HIPCHK(hipMemcpyAsync(dx, samples, (size_t)W * ld * sizeof(int16_t), hipMemcpyHostToDevice, h->stream));
    h->copy_timed = h->cfg.profile == 1;
    run_batch(h, handle_bufs(h), dx, (int)W, ld, first_sample);
}

bool dedup_on(const lss_handle* h) {
    return h->cfg.dedup < 0 ? h->sem == LSS_SEM_LSS : h->cfg.dedup != 0;
}

// Build a handle from an oss_input directory (lss_create; also each stream of a group).
std::unique_ptr<lss_handle> create_handle(const char* oss_input_dir, const lss_config* cfg, int min_rows,
                                          const int32_t* keep = nullptr, int n_keep = 0, int group_split = 0) {
    auto h = std::make_unique<lss_handle>();
    if (keep) h->keep.assign(keep, keep + n_keep);
    if (cfg) h->cfg = *cfg; else lss_config_init(&h->cfg);
    if (h->cfg.precision != LSS_PREC_FP32 && h->cfg.precision != LSS_PREC_BF16_RESCORE &&
        h->cfg.precision != LSS_PREC_BF16X3)
        throw LssError(LSS_ERR_ARG, "unknown precision mode");
    h->bf16 = h->cfg.precision == LSS_PREC_BF16_RESCORE;
    h->x3 = h->cfg.precision == LSS_PREC_BF16X3;
    if (h->bf16 && h->cfg.semantics != LSS_SEM_KILOSORT)
        throw LssError(LSS_ERR_ARG, "BF16 projection with FP32 rescoring is defined for Kilosort semantics");
    if (h->cfg.semantics != LSS_SEM_KILOSORT && h->cfg.semantics != LSS_SEM_LSS)
        throw LssError(LSS_ERR_ARG, "unknown semantics");
    if (h->cfg.chunk <= 0) throw LssError(LSS_ERR_ARG, "chunk must be positive");
    if (h->cfg.max_events <= 0) throw LssError(LSS_ERR_ARG, "max_events must be positive");
    h->dev = h->cfg.device;
    h->sem = h->cfg.semantics;
    h->max_iter = h->cfg.max_iter > 0 ? h->cfg.max_iter : (h->sem == LSS_SEM_LSS ? 50 : 100);
    if (h->max_iter > MAX_ITER) throw LssError(LSS_ERR_ARG, "max_iter too large");
    h->min_rows = min_rows;
    h->group_split = group_split;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (h->dev < 0 || h->dev >= ndev) throw LssError(LSS_ERR_ARG, "no such HIP device");
    HIPCHK(hipSetDevice(h->dev));
    // default batch: one chunk plus two margins of the longest supported template
    h->Wmax = h->cfg.max_batch > 0 ? h->cfg.max_batch : h->cfg.chunk + 2 * MAX_M;
    load(h.get(), oss_input_dir);
    h->hist.clear();
    return h;
}

}  // namespace

extern "C" {

int lss_config_init(lss_config* cfg) {
    if (!cfg) return LSS_ERR_ARG;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->device = 0;
    cfg->chunk = 3000;
    cfg->max_batch = 0;
    cfg->max_iter = 0;
    cfg->semantics = LSS_SEM_KILOSORT;
    cfg->precision = LSS_PREC_FP32;
    cfg->dedup = -1;
    cfg->max_events = 65536;
    cfg->keep_stages = 0;
    cfg->sparse = 2;
    cfg->profile = 0;
    return LSS_OK;
}

int lss_create(const char* oss_input_dir, const lss_config* cfg, lss_handle** out) {
    return guarded([&] {
        if (!oss_input_dir || !out) throw LssError(LSS_ERR_ARG, "null argument");
        *out = nullptr;
        *out = create_handle(oss_input_dir, cfg, 0).release();
    });
}

int lss_create_subset(const char* oss_input_dir, const lss_config* cfg, const int32_t* keep, int32_t n_keep,
                      lss_handle** out) {
    return guarded([&] {
        if (!oss_input_dir || !out || !keep || n_keep <= 0) throw LssError(LSS_ERR_ARG, "null argument or empty subset");
        *out = nullptr;
        *out = create_handle(oss_input_dir, cfg, 0, keep, n_keep).release();
    });
}

int lss_process_batch(lss_handle* h, const int16_t* samples, int64_t n_samples, int32_t ld,
                      const int32_t* chan_map, int64_t first_sample, lss_event* ev, int64_t ev_cap,
                      int64_t* n_ev, float* feat) {
    return guarded([&] {
        if (!h || !n_ev || ev_cap < 0) throw LssError(LSS_ERR_ARG, "null argument");
        HIPCHK(hipSetDevice(h->dev));
        auto t0 = std::chrono::steady_clock::now();
        *n_ev = 0;
        process_host_batch(h, samples, n_samples, ld, chan_map, first_sample);
        emit(h, ev, ev_cap, feat, INT64_MIN, INT64_MAX, false, n_ev);
        h->info.last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
}

int lss_process_batch_device(lss_handle* h, const int16_t* d_samples, int64_t n_samples, int32_t ld,
                             const int32_t* chan_map, int64_t first_sample, lss_event* ev, int64_t ev_cap,
                             int64_t* n_ev, float* feat) {
    return guarded([&] {
        if (!h || !n_ev || !d_samples || ev_cap < 0) throw LssError(LSS_ERR_ARG, "null argument");
        if (n_samples <= 0 || n_samples > h->Wmax)
            throw LssError(LSS_ERR_ARG, "batch of " + std::to_string(n_samples) + " samples outside 1..max_batch " +
                                            std::to_string(h->Wmax));
        HIPCHK(hipSetDevice(h->dev));
        auto t0 = std::chrono::steady_clock::now();
        *n_ev = 0;
        set_cmap(h, chan_map, ld);
        run_batch(h, handle_bufs(h), d_samples, (int)n_samples, ld, first_sample);
        emit(h, ev, ev_cap, feat, INT64_MIN, INT64_MAX, false, n_ev);
        h->info.last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
}

int lss_process_chunk(lss_handle* h, const int16_t* samples, int64_t n_samples, int32_t ld,
                      const int32_t* chan_map, int64_t first_sample, lss_event* ev, int64_t ev_cap,
                      int64_t* n_ev, float* feat) {
    return guarded([&] {
        if (!h || !n_ev || !samples || n_samples <= 0 || ev_cap < 0) throw LssError(LSS_ERR_ARG, "bad argument");
        if (ld < h->C && !chan_map) throw LssError(LSS_ERR_ARG, "row stride smaller than the channel count");
        HIPCHK(hipSetDevice(h->dev));
        auto t0 = std::chrono::steady_clock::now();
        *n_ev = 0;
        const int C = h->C, M = h->M;
        if (h->flushed) throw LssError(LSS_ERR_STATE, "stream already flushed");
        if (h->started && first_sample != h->next_sample)
            throw LssError(LSS_ERR_ARG, "chunks must be contiguous");
        // gather the chunk's channels on the host into the frame [hist | chunk]
        std::vector<int> cm(C);
        for (int c = 0; c < C; ++c) {
            cm[c] = chan_map ? chan_map[c] : h->cmap_default[c];
            if (cm[c] < 0 || cm[c] >= ld) throw LssError(LSS_ERR_ARG, "channel map entry out of range");
        }
        int64_t hist_n;
        if (!h->started) {
            // first batch: M columns repeating the first sample (KS/io.py:616-619)
            h->hist.assign((size_t)M * C, 0);
            for (int i = 0; i < M; ++i)
                for (int c = 0; c < C; ++c) h->hist[(size_t)i * C + c] = samples[cm[c]];
            hist_n = M;
        } else {
            hist_n = (int64_t)h->hist.size() / C;
        }
        const int64_t W = hist_n + n_samples;
        if (W > h->Wmax) throw LssError(LSS_ERR_ARG, "chunk longer than the handle's max_batch allows");
        h->frame.resize((size_t)W * C);
        std::memcpy(h->frame.data(), h->hist.data(), h->hist.size() * sizeof(int16_t));
        for (int64_t i = 0; i < n_samples; ++i) {
            const int16_t* src = samples + i * ld;
            int16_t* dst = &h->frame[(size_t)(hist_n + i) * C];
            for (int c = 0; c < C; ++c) dst[c] = src[cm[c]];
        }
        const int64_t col0 = first_sample - hist_n;       // absolute sample of column 0
        std::vector<int32_t> ident(C);
        std::iota(ident.begin(), ident.end(), 0);
        process_host_batch(h, h->frame.data(), W, C, ident.data(), col0);
        // accepted columns [M, W-M) -> samples [col0 + M, col0 + W - M); report events there
        const int64_t shift = -(M / 2) + h->nt0min;
        emit(h, ev, ev_cap, feat, col0 + M + shift, col0 + W - M + shift, dedup_on(h), n_ev);
        // keep the last 2M samples as the next frame's left context
        const int64_t keep = std::min<int64_t>(2 * M, W);
        h->hist.assign(h->frame.end() - keep * C, h->frame.end());
        h->started = true;
        h->next_sample = first_sample + n_samples;
        h->info.last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
}

int lss_flush(lss_handle* h, lss_event* ev, int64_t ev_cap, int64_t* n_ev, float* feat) {
    return guarded([&] {
        if (!h || !n_ev) throw LssError(LSS_ERR_ARG, "null argument");
        *n_ev = 0;
        if (!h->started) throw LssError(LSS_ERR_STATE, "nothing to flush");
        if (h->flushed) throw LssError(LSS_ERR_STATE, "already flushed");
        HIPCHK(hipSetDevice(h->dev));
        const int C = h->C, M = h->M;
        const int64_t hist_n = (int64_t)h->hist.size() / C;
        // pad the end with M copies of the last sample (KS/io.py:620-623)
        const int64_t W = hist_n + M;
        h->frame.resize((size_t)W * C);
        std::memcpy(h->frame.data(), h->hist.data(), h->hist.size() * sizeof(int16_t));
        for (int i = 0; i < M; ++i)
            std::memcpy(&h->frame[(size_t)(hist_n + i) * C], &h->hist[(size_t)(hist_n - 1) * C], C * sizeof(int16_t));
        const int64_t col0 = h->next_sample - hist_n;
        std::vector<int32_t> ident(C);
        std::iota(ident.begin(), ident.end(), 0);
        process_host_batch(h, h->frame.data(), W, C, ident.data(), col0);
        const int64_t shift = -(M / 2) + h->nt0min;
        emit(h, ev, ev_cap, feat, col0 + M + shift, col0 + W - M + shift, dedup_on(h), n_ev);
        h->flushed = true;
    });
}

int lss_query(lss_handle* h, lss_info* info) {
    return guarded([&] {
        if (!h || !info) throw LssError(LSS_ERR_ARG, "null argument");
        lss_info i = h->info;
        i.C = h->C; i.T = h->T; i.M = h->M; i.K = h->K; i.nC = h->nC;
        i.nt0min = h->nt0min; i.dt = h->dt; i.semantics = h->sem; i.Th = h->Th;
        i.max_batch = h->Wmax; i.chunk = h->cfg.chunk; i.max_iter = h->max_iter;
        i.precision = h->cfg.precision;
        if (h->d_gate_stats) {
            unsigned long long c[4];
            HIPCHK(hipSetDevice(h->dev));
            HIPCHK(hipStreamSynchronize(h->stream));
            HIPCHK(hipMemcpy(c, h->d_gate_stats, sizeof(c), hipMemcpyDeviceToHost));
            i.gate_tiles_scored = (int64_t)c[0];
            i.gate_tiles = (int64_t)c[1];
            i.gate_fallback_blocks = (int64_t)c[2];
            i.gate_blocks = (int64_t)c[3];
        }
        if (h->bf16) {
            unsigned long long c[2];
            HIPCHK(hipSetDevice(h->dev));
            HIPCHK(hipStreamSynchronize(h->stream));
            HIPCHK(hipMemcpy(c, h->d_rescnt, sizeof(c), hipMemcpyDeviceToHost));
            i.rescored = (int64_t)c[0];
            i.rescore_flips = (int64_t)c[1];
        }
        *info = i;
    });
}

int lss_debug_stage(lss_handle* h, int32_t stage, void* out, int64_t* bytes) {
    return guarded([&] {
        if (!h || !out || !bytes) throw LssError(LSS_ERR_ARG, "null argument");
        if (!h->cfg.keep_stages && stage >= 2 && stage <= 4) throw LssError(LSS_ERR_STATE, "keep_stages is off");
        if (stage == 6) {
            if (*bytes < (int64_t)PROF_WORDS * 8) throw LssError(LSS_ERR_ARG, "stage buffer too small");
            HIPCHK(hipSetDevice(h->dev));
            HIPCHK(hipMemcpy(out, h->d_prof, PROF_WORDS * 8, hipMemcpyDeviceToHost));
            *bytes = PROF_WORDS * 8;
            return;
        }
        if (stage == 7) {
            const int T = h->T, L = 2 * h->M + 1, TS = h->ctc_rows;
            const int64_t need = (int64_t)T * T * L * 4;
            if (*bytes < need) throw LssError(LSS_ERR_ARG, "stage buffer too small");
            HIPCHK(hipSetDevice(h->dev));
            std::vector<float> ci((size_t)T * TS * LP);
            HIPCHK(hipMemcpy(ci.data(), h->d_ctc, ci.size() * 4, hipMemcpyDeviceToHost));
            float* of = static_cast<float*>(out);
            for (int sp = 0; sp < T; ++sp)
                for (int tp = 0; tp < T; ++tp)
                    std::memcpy(of + ((size_t)h->orig[tp] * T + h->orig[sp]) * L, &ci[((size_t)sp * TS + tp) * LP],
                                (size_t)L * 4);
            *bytes = need;
            return;
        }
        if (h->last_W <= 0) throw LssError(LSS_ERR_STATE, "no batch processed yet");
        HIPCHK(hipSetDevice(h->dev));
        HIPCHK(hipStreamSynchronize(h->stream));
        const int W = h->last_W, C = h->C, T = h->T, K = h->K;
        int64_t need = 0;
        if (stage >= 8 && stage <= 10) {
            if (!h->pre) throw LssError(LSS_ERR_STATE, "preprocessing is off");
            if (stage == 8) {
                if (*bytes < (int64_t)C * W * 4) throw LssError(LSS_ERR_ARG, "stage buffer too small");
                HIPCHK(hipMemcpy2D(out, (size_t)W * 4, (h->pre & PRE_HP) ? h->d_Xp : h->d_X1, (size_t)h->PS * 4,
                                   (size_t)W * 4, C, hipMemcpyDeviceToHost));
                *bytes = (int64_t)C * W * 4;
            } else if (stage == 9) {
                if (*bytes < (int64_t)C * 4) throw LssError(LSS_ERR_ARG, "stage buffer too small");
                HIPCHK(hipMemcpy(out, h->d_pmean, (size_t)C * 4, hipMemcpyDeviceToHost));
                *bytes = (int64_t)C * 4;
            } else {
                if (*bytes < (int64_t)W * 4) throw LssError(LSS_ERR_ARG, "stage buffer too small");
                HIPCHK(hipMemcpy(out, h->d_pmed, (size_t)W * 4, hipMemcpyDeviceToHost));
                *bytes = (int64_t)W * 4;
            }
            return;
        }
        if (stage == 0) need = (int64_t)C * W * 4;
        else if (stage == 1) need = (int64_t)K * C * W * 4;
        else if (stage == 2) need = (int64_t)T * W * 4;
        else if (stage == 3 || stage == 4) need = (int64_t)W * 4;
        else if (stage == 5) need = (int64_t)T * 4;
        else if (stage == 6) need = (int64_t)PROF_WORDS * 8;
        else throw LssError(LSS_ERR_ARG, "unknown stage");
        if (*bytes < need) throw LssError(LSS_ERR_ARG, "stage buffer too small");
        char* o = static_cast<char*>(out);
        if (stage == 0) {
            HIPCHK(hipMemcpy2D(o, (size_t)W * 4, h->d_Xw + XOFF, (size_t)h->XS * 4, (size_t)W * 4, C,
                               hipMemcpyDeviceToHost));
        } else if (stage == 1) {
            for (int k = 0; k < K; ++k)
                HIPCHK(hipMemcpy2D(o + (size_t)k * C * W * 4, (size_t)W * 4, h->d_Bpc + (size_t)k * h->CP * h->BS,
                                   (size_t)h->BS * 4, (size_t)W * 4, C, hipMemcpyDeviceToHost));
        } else if (stage == 2) {
            // rows back in file template order
            std::vector<float> tmp((size_t)T * W);
            HIPCHK(hipMemcpy2D(tmp.data(), (size_t)W * 4, h->d_dbgB, (size_t)h->BS * 4, (size_t)W * 4, T,
                               hipMemcpyDeviceToHost));
            float* of = reinterpret_cast<float*>(o);
            for (int tp = 0; tp < T; ++tp)
                std::memcpy(of + (size_t)h->orig[tp] * W, &tmp[(size_t)tp * W], (size_t)W * 4);
        } else if (stage == 3) {
            HIPCHK(hipMemcpy(o, h->d_dbgCf, (size_t)W * 4, hipMemcpyDeviceToHost));
        } else if (stage == 4) {
            std::vector<int> tmp(W);
            HIPCHK(hipMemcpy(tmp.data(), h->d_dbgIm, (size_t)W * 4, hipMemcpyDeviceToHost));
            int* oi = reinterpret_cast<int*>(o);
            for (int w = 0; w < W; ++w) oi[w] = h->orig[tmp[w]];
        } else if (stage == 5) {
            std::vector<float> tmp(T);
            HIPCHK(hipMemcpy(tmp.data(), h->d_nm, (size_t)T * 4, hipMemcpyDeviceToHost));
            float* of = reinterpret_cast<float*>(o);
            for (int tp = 0; tp < T; ++tp) of[h->orig[tp]] = tmp[tp];
        }
        *bytes = need;
    });
}

int lss_autotune(lss_handle* h, const int16_t* d_samples, int64_t n, int32_t ld, int32_t reps, lss_tune_result* out) {
    return guarded([&] {
        if (!h || !d_samples || !out) throw LssError(LSS_ERR_ARG, "null argument");
        if (n <= 0 || n > h->Wmax) throw LssError(LSS_ERR_ARG, "batch width outside 1..max_batch");
        if (reps < 1 || reps > 1000) throw LssError(LSS_ERR_ARG, "reps outside 1..1000");
        HIPCHK(hipSetDevice(h->dev));
        set_cmap(h, nullptr, ld);
        const int W = (int)n;
        const lss_info saved = h->info;
        const int32_t prof = h->cfg.profile;   // restored afterwards (the timings are host-side)
        h->cfg.profile = 0;
        *out = lss_tune_result{};
        int cand_a[4];
        const std::vector<int> cand(cand_a, cand_a + split_candidates(h, W, cand_a));
        int best = 0;
        try {
            for (size_t i = 0; i < cand.size(); ++i) {
                h->force_split = cand[i];
                int CW = 0, R = 0, S = 1;
                peel_geometry(h, W, CW, R, S);
                std::vector<double> t;
                for (int r = 0; r <= reps; ++r) {
                    const auto t0 = std::chrono::steady_clock::now();
                    run_batch(h, handle_bufs(h), d_samples, W, ld, 0);
                    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                    if (r > 0) t.push_back(ms);
                }
                std::sort(t.begin(), t.end());
                out->split[i] = S;
                out->rows[i] = R;
                out->ms[i] = t[t.size() / 2];
                if (out->ms[i] < out->ms[best]) best = (int)i;
            }
        } catch (...) {
            h->force_split = 0;
            h->cfg.profile = prof;
            h->info = saved;
            throw;
        }
        h->force_split = 0;
        h->cfg.profile = prof;
        h->info = saved;
        out->n = (int32_t)cand.size();
        out->chosen = best;
        h->tuned_split[W] = out->split[best];
    });
}

int lss_last_positions(lss_handle* h, float* xy, int64_t cap, int64_t* n) {
    return guarded([&] {
        if (!h || !n) throw LssError(LSS_ERR_ARG, "null argument");
        if (!h->pos) throw LssError(LSS_ERR_STATE, "no probe geometry (xc.npy / yc.npy) was loaded");
        const int64_t m = (int64_t)h->last_pos.size() / 2;
        *n = m;
        if (m > cap) throw LssError(LSS_ERR_CAPACITY, "position buffer too small");
        if (m && !xy) throw LssError(LSS_ERR_ARG, "null argument");
        if (m) std::memcpy(xy, h->last_pos.data(), sizeof(float) * 2 * m);
    });
}

int lss_stage_times(lss_handle* h, double* ms, int64_t* batches, int32_t reset) {
    return guarded([&] {
        if (!h || !ms || !batches) throw LssError(LSS_ERR_ARG, "null argument");
        if (!h->cfg.profile) throw LssError(LSS_ERR_STATE, "profile is off");
        for (int i = 0; i < LSS_NSTAGES; ++i) ms[i] = h->stage_ms[i];
        *batches = h->timed_batches;
        if (reset) {
            for (double& v : h->stage_ms) v = 0.0;
            h->timed_batches = 0;
        }
    });
}

int lss_destroy(lss_handle* h) {
    return guarded([&] {
        if (!h) return;
        delete h;
    });
}

const char* lss_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

// ---------------------------------------------------------------------------
// Groups: several probe streams matched by one launch sequence per tick,
// captured once into a hipGraph (north_star form (2)).  Each stream keeps its
// own tables and buffers (an internal handle); the kernels are the batched
// forms of the single-stream ones, the peel partitions its workgroups among
// the streams (a barrier counter per stream), and every table is padded to the
// group's register tile height so one kernel instance serves all template counts.
// ---------------------------------------------------------------------------
// A launch lane: its own HIP stream, captured graph and input-pointer table.
// A lane's own per-batch buffers for one stream, so that a pipelined group can
// have a stream in flight on two lanes at once: the device buffers laid out as
// the handle's, the events and features in mapped host memory the peel writes
// into, the proposal network's outputs and the gate's tile mask.
struct LaneBufs {
    unsigned char* dev = nullptr;
    void* host = nullptr;
    BatchBufs b;
    OutHeader* hdr = nullptr;            // host views of the outputs
    lss_event* ev = nullptr;
    float* feat = nullptr;
    float* pos = nullptr;
    PnBuffers pnb{};                     // the handle's, with this lane's z and p
    float* pz = nullptr;
    float* pp = nullptr;
    void release() {
        if (dev) (void)hipFree(dev);
        if (host) (void)hipHostFree(host);
        if (pz) (void)hipFree(pz);
        if (pp) (void)hipFree(pp);
        if (b.gate) (void)hipFree(b.gate);
        dev = nullptr; host = nullptr; pz = pp = nullptr; b.gate = nullptr;
    }
};

// Allocate a lane's buffers for a handle (all zero: no tag of any epoch).
static void alloc_lane_bufs(const lss_handle* h, LaneBufs& lb) {
    const size_t tag_bytes = 256 + ((size_t)h->ptag_words * sizeof(unsigned long long) + 255) / 256 * 256;
    size_t need = 0;
    (void)layout_bufs(h, nullptr, tag_bytes, &need);
    HIPCHK(hipMalloc(reinterpret_cast<void**>(&lb.dev), need));
    HIPCHK(hipMemset(lb.dev, 0, need));
    lb.b = layout_bufs(h, lb.dev, tag_bytes, &need);
    // the outputs: mapped, coherent host memory, as the handle's
    const size_t fw = (size_t)h->nC * h->K;
    const size_t out_bytes = sizeof(OutHeader) + sizeof(lss_event) * (size_t)h->ev_cap +
                             sizeof(float) * (size_t)h->ev_cap * fw + sizeof(float) * 2 * (size_t)h->ev_cap;
    HIPCHK(hipHostMalloc(&lb.host, out_bytes, hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(lb.host, 0, out_bytes);
    void* dout = nullptr;
    HIPCHK(hipHostGetDevicePointer(&dout, lb.host, 0));
    lb.hdr = static_cast<OutHeader*>(lb.host);
    lb.ev = reinterpret_cast<lss_event*>(lb.hdr + 1);
    lb.feat = reinterpret_cast<float*>(lb.ev + h->ev_cap);
    lb.b.hdr = static_cast<OutHeader*>(dout);
    lb.b.ev = reinterpret_cast<lss_event*>(lb.b.hdr + 1);
    lb.b.feat = reinterpret_cast<float*>(lb.b.ev + h->ev_cap);
    lb.b.pos = nullptr;
    if (h->d_xc) {
        lb.pos = lb.feat + (size_t)h->ev_cap * fw;
        lb.b.pos = lb.b.feat + (size_t)h->ev_cap * fw;
    }
    if (h->cfg.keep_stages) { lb.b.dbgB = h->d_dbgB; lb.b.dbgCf = h->d_dbgCf; lb.b.dbgIm = h->d_dbgIm; }
    if (h->prop) {
        lb.pnb = h->pnb;
        const size_t zp = (size_t)h->pnb.G * h->pnb.Wp;
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&lb.pz), zp * sizeof(float)));
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&lb.pp), zp * sizeof(float)));
        lb.pnb.z = lb.pz;
        lb.pnb.p = lb.pp;
        const size_t nm = (size_t)((h->Wmax + 63) / 64) * (h->TP / GEMM_BM);
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&lb.b.gate), nm));
        HIPCHK(hipMemset(lb.b.gate, 0, nm));
    }
}

// Lanes run concurrently on disjoint sets of streams (a stream idle in a
// lane's launch is untouched by it).  In a pipelined group lanes 1.. have
// their own per-batch buffers, so consecutive batches of the same streams can
// be in flight on two lanes: the next batch's front stages run beside the
// current peel, and the peels run one after another.
struct GroupLane {
    int id = 0;
    std::vector<LaneBufs> own;           // pipelined groups, lanes 1..: per stream
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph2 = nullptr;         // cfg.profile: the peel in a graph of its own
    hipGraphExec_t exec2 = nullptr;
    const int16_t** h_xptr = nullptr;    // pinned: this launch's input pointers (null = idle)
    const int16_t** d_xptr = nullptr;    // device copy the kernels read
    // host input: per stream, a pinned staging copy of the caller's batch, so the
    // copy to the device is a DMA from pinned memory (reused after the lane's sync)
    std::vector<int16_t*> h_stage;
    std::vector<size_t> h_stage_elems;
    std::vector<int32_t> ld_captured;
    WhitenBatch wb{};
    PcconvBatch cb{};
    ProjectBatch pb{};
    PeelBatch kb{};
    ResetBatch rb{};
    // cfg.profile: the tick is two graphs, the front stages and the peel, with events
    // recorded on the lane's stream before, between and after them; the device time
    // they measured on this lane (front stages, peel)
    hipEvent_t ev[3] = {};
    double ms[2] = {};
    int64_t timed = 0;
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (exec2) (void)hipGraphExecDestroy(exec2);
        if (graph2) (void)hipGraphDestroy(graph2);
        exec = exec2 = nullptr;
        graph = graph2 = nullptr;
    }
};
constexpr int GROUP_LANES = 4;

struct lss_group {
    std::vector<std::unique_ptr<lss_handle>> hs;
    int n = 0, W = 0, R = 0, CW = 0, dev = 0;
    bool profile = false;
    // pipelined: every lane's peel waits for the peel enqueued before it (one peel grid
    // co-resident at a time); the order of enqueueing is taken under `seq`
    bool pipelined = false;
    hipEvent_t peel_ev = nullptr;
    std::mutex seq;
    std::vector<int> S;   // per stream: template subsets per column block in the peel grid
    GroupLane lanes[GROUP_LANES];
    // Lanes run on several host threads.  Allocations, synchronous copies and graph
    // captures go under this lock, so no lane's capture sees another thread's
    // allocating or synchronizing call in flight.
    std::mutex mu;
    ~lss_group() {
        (void)hipSetDevice(dev);
        for (auto& L : lanes) {
            L.drop_graph();
            for (auto& e : L.ev)
                if (e) (void)hipEventDestroy(e);
            if (L.h_xptr) (void)hipHostFree(L.h_xptr);
            for (int16_t* b : L.h_stage)
                if (b) (void)hipHostFree(b);
            if (L.d_xptr) (void)hipFree(L.d_xptr);
            if (L.stream) (void)hipStreamDestroy(L.stream);
            for (auto& lb : L.own) lb.release();
        }
        if (peel_ev) (void)hipEventDestroy(peel_ev);
    }
};

namespace {

int64_t peek_templates(const std::string& dir_in) {
    std::string dir = dir_in;
    if (!dir.empty() && dir.back() != '/') dir += '/';
    auto a = npy_load(dir + "Wall3.npy");
    if (a.shape.size() != 3) throw LssError(LSS_ERR_IO, "Wall3.npy must be [T][K][C]");
    return a.shape[0];
}

// (Re)build the job tables and capture the tick's launch sequence into a graph.
void group_capture(lss_group* g, GroupLane& L, const int32_t* ld) {
    const int n = g->n;
    L.wb = WhitenBatch{}; L.cb = PcconvBatch{}; L.pb = ProjectBatch{}; L.kb = PeelBatch{}; L.rb = ResetBatch{};
    L.wb.n = L.cb.n = L.pb.n = L.kb.n = L.rb.n = n;
    L.kb.xcd = n > 1;                   // the peel's streams on their own XCDs
    PreBatch pre{};
    // lanes 1.. of a pipelined group work in buffers of their own
    const bool mine = g->pipelined && L.id > 0;
    if (mine && (int)L.own.size() < n) {
        L.own.resize(n);
        for (int s2 = 0; s2 < n; ++s2) alloc_lane_bufs(g->hs[s2].get(), L.own[s2]);
    }
    for (int s2 = 0; s2 < n; ++s2) {
        lss_handle* h = g->hs[s2].get();
        const BatchBufs bb = mine ? L.own[s2].b : handle_bufs(h);
        BatchPlan bp;
        plan_batch(h, bb, nullptr, &L.d_xptr[s2], g->W, ld[s2], g->CW, g->R, g->S[s2], bp);
        if (bp.has_pre) pre.job[pre.n++] = bp.pre;
        L.wb.job[s2] = bp.wj;
        L.wb.start[s2 + 1] = L.wb.start[s2] + gemm_blocks(bp.wj.mtiles, bp.wj.ntiles);
        L.cb.job[s2] = bp.cj;
        L.cb.start[s2 + 1] = L.cb.start[s2] + pc_job_blocks(bp.cj);
        L.pb.job[s2] = bp.pj;
        L.pb.start[s2 + 1] = L.pb.start[s2] + gemm_blocks(bp.pj.mtiles, bp.pj.ntiles);
        L.kb.p[s2] = bp.pp;
        L.kb.g_start[s2 + 1] = L.kb.g_start[s2] + bp.G;
        L.rb.bar[s2] = bb.bar; L.rb.iter_cnt[s2] = bb.iter_cnt; L.rb.n_iter[s2] = h->sync_words;
        L.rb.err[s2] = bb.err; L.rb.prof[s2] = bb.prof; L.rb.active[s2] = &L.d_xptr[s2];
        L.rb.epoch[s2] = bb.epoch; L.rb.ptag[s2] = bb.ptag; L.rb.ptag_words[s2] = h->ptag_words;
    }
    // GEMM launches of 2, 4 or 8 streams: each stream's tiles on its own share of the
    // XCDs, so an XCD's L2 keeps the template and whitening operands of 8 / n streams
    if (n > 1 && 8 % n == 0) {
        int wmax = 0, pmax = 0, cmax = 0;
        for (int s2 = 0; s2 < n; ++s2) {
            wmax = std::max(wmax, gemm_blocks_ilv(L.wb.job[s2].mtiles, L.wb.job[s2].ntiles, n));
            pmax = std::max(pmax, gemm_blocks_ilv(L.pb.job[s2].mtiles, L.pb.job[s2].ntiles, n));
            cmax = std::max(cmax, L.cb.start[s2 + 1] - L.cb.start[s2]);
        }
        L.wb.ilv = L.pb.ilv = L.cb.ilv = 1;
        L.wb.start[n] = n * wmax;
        L.pb.start[n] = n * pmax;
        L.cb.start[n] = n * cmax;
    }
    L.drop_graph();
    hipStream_t st = L.stream;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    launch_preprocess(pre, st);
    launch_whiten(L.wb, &L.rb, st);
    {   // the proposal networks of all gated streams: one fused dispatch, then the gate.
        // (On a branch of the graph beside the PC conv they ran slower than one after the
        // other: the PC conv's small workgroups hold the CUs the network's need.)
        PnFwdJob fj[PN_MAX_JOBS];
        PnGateJob gj[PN_MAX_JOBS];
        int ng = 0;
        for (int s2 = 0; s2 < n; ++s2) {
            lss_handle* h = g->hs[s2].get();
            if (!h->prop) continue;
            const PnBuffers& pnb = mine ? L.own[s2].pnb : h->pnb;
            PnGate gt = h->gate;
            if (mine) gt.mask = L.own[s2].b.gate;
            fj[ng] = PnFwdJob{mine ? L.own[s2].b.Xw : h->d_Xw, h->XS, XOFF, h->C, g->W, h->prop->wt, pnb, &L.d_xptr[s2]};
            gj[ng] = PnGateJob{pnb, gt, g->W, &L.d_xptr[s2]};
            ++ng;
        }
        pn_infer_batch(fj, ng, st);
        pn_gate_batch(gj, ng, st);
        launch_pcconv(L.cb, g->hs[0]->K, st);
    }
    launch_project(L.pb, g->hs[0]->cfg.precision, st);
    // profiled and pipelined groups: the front stages and the peel are two graphs
    const bool split = g->profile || g->pipelined;
    if (split) {
        HIPCHK(hipStreamEndCapture(st, &L.graph));
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    }
    HIPCHK(launch_peel(L.kb, g->R, false, st));
    if (split) {
        HIPCHK(hipStreamEndCapture(st, &L.graph2));
        HIPCHK(hipGraphInstantiate(&L.exec2, L.graph2, nullptr, nullptr, 0));
    } else {
        HIPCHK(hipStreamEndCapture(st, &L.graph));
    }
    HIPCHK(hipGraphInstantiate(&L.exec, L.graph, nullptr, nullptr, 0));
    L.ld_captured.assign(ld, ld + n);
}

}  // namespace

extern "C" {

int lss_group_create(int32_t n_streams, const char* const* oss_input_dirs, const lss_config* cfg, lss_group** out) {
    return guarded([&] {
        if (!out || !oss_input_dirs || !cfg) throw LssError(LSS_ERR_ARG, "null argument");
        *out = nullptr;
        if (n_streams < 1 || n_streams > MAX_JOBS)
            throw LssError(LSS_ERR_ARG, "a group holds 1.." + std::to_string(MAX_JOBS) + " streams");
        if (cfg->max_batch <= 0) throw LssError(LSS_ERR_ARG, "a group needs max_batch (the framed batch width)");
        auto g = std::make_unique<lss_group>();
        g->n = n_streams;
        g->W = cfg->max_batch;
        g->dev = cfg->device;
        int64_t tmax = 0;
        std::vector<int64_t> ts(n_streams);
        for (int s2 = 0; s2 < n_streams; ++s2) {
            if (!oss_input_dirs[s2]) throw LssError(LSS_ERR_ARG, "null directory");
            ts[s2] = peek_templates(oss_input_dirs[s2]);
            tmax = std::max(tmax, ts[s2]);
        }
        HIPCHK(hipSetDevice(g->dev));
        // One peel kernel instance for all streams: the register tile of the largest T.
        // With Kilosort rules a shorter tile may serve instead, each stream's templates
        // split over ceil(T / tile rows) workgroups per column block, when that grid is
        // co-resident: 16 rows per thread first, then 32.  peel_split 1 keeps one subset;
        // n > 1 asks for the tile that splits the largest stream n ways, when it fits.
        g->CW = MIN_CW;
        g->R = peel_rows((int)tmax, g->CW);
        const int NB = (g->W + MIN_CW - 1) / MIN_CW;
        auto subsets = [](int R, int64_t T) { return (int)((T + REG_NRG * R - 1) / (REG_NRG * R)); };
        auto split_ok = [&](int R) {
            if (R <= 0 || R >= g->R || subsets(R, tmax) > 16) return false;
            int64_t sum = 0;
            for (int64_t T : ts) sum += (int64_t)subsets(R, T) * NB;
            return sum <= peel_max_blocks(g->dev, R);
        };
        int Rs = 0;
        const int want = cfg->peel_split;
        if (g->R > 0 && tmax <= REG_MAX_T && cfg->semantics == LSS_SEM_KILOSORT && want != 1) {
            if (want > 1 && split_ok(split_rows((int)tmax, want))) Rs = split_rows((int)tmax, want);
            for (int r : {16, 32})
                if (Rs == 0 && split_ok(r) && (want <= 1 || subsets(r, tmax) <= want)) Rs = r;
        }
        if (Rs > 0) g->R = Rs;
        g->S.assign(n_streams, 1);
        for (int s2 = 0; s2 < n_streams; ++s2)
            if (Rs > 0) g->S[s2] = subsets(Rs, ts[s2]);
        const int rows = g->R > 0 ? REG_NRG * g->R : 0;
        for (int s2 = 0; s2 < n_streams; ++s2) {
            g->hs.push_back(create_handle(oss_input_dirs[s2], cfg, rows, nullptr, 0, g->S[s2]));
            if (g->hs.back()->K != g->hs[0]->K || g->hs.back()->M != g->hs[0]->M)
                throw LssError(LSS_ERR_ARG, "streams of a group must share K and M");
        }
        if (g->R == 0) g->CW = 2 * MIN_CW;
        int gsum = 0;
        for (size_t s2 = 0; s2 < g->hs.size(); ++s2) {
            lss_handle* h = g->hs[s2].get();
            const int G = g->S[s2] * ((g->W + g->CW - 1) / g->CW);
            gsum += G;
            h->info.peel_cw = g->CW; h->info.peel_rows = g->R; h->info.peel_split = g->S[s2];
            h->info.peel_groups = G;
        }
        if (gsum > g->hs[0]->peel_cap[peel_cap_index(g->R)])
            throw LssError(LSS_ERR_ARG, "the group's peel grid (" + std::to_string(gsum) +
                                            " workgroups) is not co-resident on this device");
        HIPCHK(hipSetDevice(g->dev));
        g->profile = cfg->profile != 0;
        for (int li = 0; li < GROUP_LANES; ++li) g->lanes[li].id = li;
        for (auto& L : g->lanes) {
            if (g->profile)
                for (auto& e : L.ev) HIPCHK(hipEventCreate(&e));
            HIPCHK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
            HIPCHK(hipHostMalloc((void**)&L.h_xptr, sizeof(void*) * MAX_JOBS, hipHostMallocDefault));
            HIPCHK(hipMalloc((void**)&L.d_xptr, sizeof(void*) * MAX_JOBS));
        }
        *out = g.release();
    });
}

int lss_group_process(lss_group* g, const int16_t* const* samples, const int32_t* ld, int32_t device_input,
                      const int64_t* first_sample, lss_event* const* ev, const int64_t* ev_cap, int64_t* n_ev,
                      float* const* feat) {
    return lss_group_process_lane(g, 0, samples, ld, device_input, first_sample, ev, ev_cap, n_ev, feat);
}

int lss_group_process_lane(lss_group* g, int32_t lane, const int16_t* const* samples, const int32_t* ld,
                           int32_t device_input, const int64_t* first_sample, lss_event* const* ev,
                           const int64_t* ev_cap, int64_t* n_ev, float* const* feat) {
    return guarded([&] {
        if (!g || !samples || !ld || !first_sample || !ev || !ev_cap || !n_ev)
            throw LssError(LSS_ERR_ARG, "null argument");
        if (lane < 0 || lane >= GROUP_LANES) throw LssError(LSS_ERR_ARG, "no such lane");
        if (g->pipelined && !device_input)
            throw LssError(LSS_ERR_ARG, "a pipelined group takes device-resident batches");
        GroupLane& L = g->lanes[lane];
        HIPCHK(hipSetDevice(g->dev));
        auto t0 = std::chrono::steady_clock::now();
        // a call that failed after queueing its copies may have left them in flight
        HIPCHK(hipStreamSynchronize(L.stream));
        const int n = g->n;
        int active = 0;
        for (int s2 = 0; s2 < n; ++s2) {
            lss_handle* h = g->hs[s2].get();
            n_ev[s2] = 0;
            if (!samples[s2]) {             // idle this tick: every kernel of the stream returns at once
                L.h_xptr[s2] = nullptr;
                continue;
            }
            ++active;
            int16_t* dx = nullptr;
            const size_t elems = (size_t)g->W * ld[s2];
            {
                std::lock_guard<std::mutex> lk(g->mu);
                set_cmap(h, nullptr, ld[s2]);
                if (!device_input) {
                    dx = ensure_dx(h, elems);
                    if (L.h_stage.size() < (size_t)n) {
                        L.h_stage.resize(n, nullptr);
                        L.h_stage_elems.resize(n, 0);
                    }
                    if (L.h_stage_elems[s2] < elems) {
                        if (L.h_stage[s2]) HIPCHK(hipHostFree(L.h_stage[s2]));
                        L.h_stage[s2] = nullptr;
                        L.h_stage_elems[s2] = 0;
                        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&L.h_stage[s2]), elems * sizeof(int16_t),
                                             hipHostMallocDefault));
                        L.h_stage_elems[s2] = elems;
                    }
                }
            }
            if (device_input) {
                L.h_xptr[s2] = samples[s2];
            } else {
                // the lane's previous launch has synchronized, so its staging copy is free
                std::memcpy(L.h_stage[s2], samples[s2], elems * sizeof(int16_t));
                HIPCHK(hipMemcpyAsync(dx, L.h_stage[s2], elems * sizeof(int16_t), hipMemcpyHostToDevice, L.stream));
                L.h_xptr[s2] = dx;
            }
        }
        if (!active) return;
        if (!L.exec || !std::equal(ld, ld + n, L.ld_captured.begin())) {
            std::lock_guard<std::mutex> lk(g->mu);
            group_capture(g, L, ld);
        }
        // this launch's input pointers (null: idle stream), read by every kernel of the
        // graph; copied ahead of it on the lane's stream rather than as a graph node
        HIPCHK(hipMemcpyAsync(L.d_xptr, L.h_xptr, sizeof(void*) * n, hipMemcpyHostToDevice, L.stream));
        if (g->pipelined) {
            // the peel waits for the peel enqueued before it on any lane, so one peel grid
            // is co-resident at a time; the front stages run beside the other lane's peel
            std::lock_guard<std::mutex> lk(g->seq);
            if (g->profile) HIPCHK(hipEventRecord(L.ev[0], L.stream));
            HIPCHK(hipGraphLaunch(L.exec, L.stream));
            if (g->profile) HIPCHK(hipEventRecord(L.ev[1], L.stream));
            HIPCHK(hipStreamWaitEvent(L.stream, g->peel_ev, 0));
            HIPCHK(hipGraphLaunch(L.exec2, L.stream));
            HIPCHK(hipEventRecord(g->peel_ev, L.stream));
            if (g->profile) HIPCHK(hipEventRecord(L.ev[2], L.stream));
        } else if (g->profile) {
            HIPCHK(hipEventRecord(L.ev[0], L.stream));
            HIPCHK(hipGraphLaunch(L.exec, L.stream));
            HIPCHK(hipEventRecord(L.ev[1], L.stream));
            HIPCHK(hipGraphLaunch(L.exec2, L.stream));
            HIPCHK(hipEventRecord(L.ev[2], L.stream));
        } else {
            HIPCHK(hipGraphLaunch(L.exec, L.stream));
        }
        HIPCHK(hipStreamSynchronize(L.stream));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (g->profile) {
            float a = 0.f, b = 0.f;
            HIPCHK(hipEventElapsedTime(&a, L.ev[0], L.ev[1]));
            HIPCHK(hipEventElapsedTime(&b, L.ev[1], L.ev[2]));
            L.ms[0] += a;
            L.ms[1] += b;
            L.timed += 1;
        }
        // every stream's results are checked and returned; a failure of one stream
        // does not hide the others' events.  The call then reports the failed
        // streams (their n_ev as emit left them), a timeout taking precedence.
        std::string failed;
        int code = LSS_OK;
        // pipelined lanes may emit for the same handles at once: the handles' host state
        // (event counts, last batch) is updated under the group lock
        std::unique_lock<std::mutex> lk(g->mu, std::defer_lock);
        if (g->pipelined) lk.lock();
        const bool mine = g->pipelined && L.id > 0;
        for (int s2 = 0; s2 < n; ++s2) {
            if (!samples[s2]) continue;
            lss_handle* h = g->hs[s2].get();
            h->cur_first = first_sample[s2];
            h->info.last_ms = ms;
            try {
                if (mine) {
                    const LaneBufs& lb = L.own[s2];
                    check_peel(h, *lb.hdr, g->W);
                    emit_from(h, *lb.hdr, lb.ev, lb.feat, lb.pos, ev[s2], ev_cap[s2], feat ? feat[s2] : nullptr,
                              INT64_MIN, INT64_MAX, false, &n_ev[s2]);
                } else {
                    check_peel(h, *h->hdr, g->W);
                    emit(h, ev[s2], ev_cap[s2], feat ? feat[s2] : nullptr, INT64_MIN, INT64_MAX, false, &n_ev[s2]);
                }
            } catch (const LssError& e) {
                failed += " stream " + std::to_string(s2) + ": " + e.what() + ";";
                if (code != LSS_ERR_TIMEOUT) code = e.code;
            }
        }
        if (code != LSS_OK) throw LssError(code, "group launch:" + failed);
    });
}

int lss_group_stage_times(lss_group* g, int32_t lane, double* ms, int64_t* launches, int32_t reset) {
    return guarded([&] {
        if (!g || !ms || !launches) throw LssError(LSS_ERR_ARG, "null argument");
        if (lane < 0 || lane >= GROUP_LANES) throw LssError(LSS_ERR_ARG, "no such lane");
        if (!g->profile) throw LssError(LSS_ERR_STATE, "the group was created without profile");
        GroupLane& L = g->lanes[lane];
        ms[0] = L.ms[0];
        ms[1] = L.ms[1];
        *launches = L.timed;
        if (reset) {
            L.ms[0] = L.ms[1] = 0.0;
            L.timed = 0;
        }
    });
}

int lss_group_debug_stage(lss_group* g, int32_t stream, int32_t stage, void* out, int64_t* bytes) {
    return guarded([&] {
        if (!g || stream < 0 || stream >= g->n) throw LssError(LSS_ERR_ARG, "bad argument");
        if (stage != 6) throw LssError(LSS_ERR_ARG, "groups report stage 6 (the peel's phase clock) only");
        const int rc = lss_debug_stage(g->hs[stream].get(), stage, out, bytes);
        if (rc != LSS_OK) throw LssError(rc, g_last_error);
    });
}

int lss_group_query(lss_group* g, int32_t stream, lss_info* info) {
    return guarded([&] {
        if (!g || !info || stream < 0 || stream >= g->n) throw LssError(LSS_ERR_ARG, "bad argument");
        const int rc = lss_query(g->hs[stream].get(), info);
        if (rc != LSS_OK) throw LssError(rc, g_last_error);
    });
}

int lss_group_set_proposal(lss_group* g, int32_t stream, lss_proposal* net, float lo, float hi, int32_t reach) {
    return guarded([&] {
        if (!g || stream < 0 || stream >= g->n) throw LssError(LSS_ERR_ARG, "bad argument");
        HIPCHK(hipSetDevice(g->dev));
        for (auto& L : g->lanes) HIPCHK(hipStreamSynchronize(L.stream));
        const int rc = lss_set_proposal(g->hs[stream].get(), net, lo, hi, reach);
        if (rc != LSS_OK) throw LssError(rc, g_last_error);
        for (auto& L : g->lanes) L.drop_graph();
    });
}

int lss_group_set_pipelined(lss_group* g, int32_t on) {
    return guarded([&] {
        if (!g) throw LssError(LSS_ERR_ARG, "null argument");
        HIPCHK(hipSetDevice(g->dev));
        std::lock_guard<std::mutex> lk(g->mu);
        for (auto& L : g->lanes) HIPCHK(hipStreamSynchronize(L.stream));
        if (on && !g->peel_ev) HIPCHK(hipEventCreateWithFlags(&g->peel_ev, hipEventDisableTiming));
        g->pipelined = on != 0;
        for (auto& L : g->lanes) L.drop_graph();    // recaptured on the next launch (split graphs, own buffers)
    });
}

int lss_group_destroy(lss_group* g) {
    return guarded([&] { delete g; });
}

// ---- spike-proposal network ----

int32_t lss_proposal_n_params(void) { return PN_NPARAMS; }

int lss_proposal_create(int32_t device, int32_t channels, int32_t max_batch, uint64_t seed, lss_proposal** out) {
    return guarded([&] {
        if (!out || channels <= 0 || max_batch <= 0) throw LssError(LSS_ERR_ARG, "bad argument");
        if (pn_groups(channels) > 64) throw LssError(LSS_ERR_ARG, "at most 1024 channels");
        *out = nullptr;
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw LssError(LSS_ERR_ARG, "no such HIP device");
        HIPCHK(hipSetDevice(device));
        auto net = std::make_unique<lss_proposal>();
        net->dev = device;
        net->C = channels;
        net->G = pn_groups(channels);
        net->max_W = max_batch;
        HIPCHK(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking));
        net->wt.params = dalloc<float>(PN_NPARAMS);
        net->wt.wf1 = dalloc<__bf16>(PN_CH * PN_KP1);
        net->wt.wf2 = dalloc<__bf16>(PN_CH * PN_KP);
        net->wt.wf3 = dalloc<__bf16>(PN_CH * PN_KP);
        net->wt.wb2 = dalloc<__bf16>(PN_CH * PN_KP);
        net->wt.wb3 = dalloc<__bf16>(PN_CH * PN_KP);
        net->wt.wt1 = dalloc<__bf16>(PN_CH * PN_KT1);
        net->wt.wt2 = dalloc<__bf16>(PN_CH * PN_KP);
        net->wt.wt3 = dalloc<__bf16>(PN_CH * PN_KP);
        net->d_m = dalloc<float>(PN_NPARAMS);
        net->d_v = dalloc<float>(PN_NPARAMS);
        net->d_grad = dalloc<float>(PN_NPARAMS);
        net->d_xw = dalloc<float>((size_t)channels * max_batch);
        alloc_pn_buffers(net->buf, net->G, max_batch);
        net->d_y = dalloc<uint8_t>((size_t)net->G * net->buf.Wp);
        // He initialisation (normal, std sqrt(2 / fan_in)), zero biases
        std::vector<float> w(PN_NPARAMS, 0.f);
        std::mt19937_64 rng(seed);
        std::normal_distribution<float> nd(0.f, 1.f);
        auto fill = [&](int off, int n, int fan_in) {
            const float sd = std::sqrt(2.f / (float)fan_in);
            for (int i = 0; i < n; ++i) w[off + i] = sd * nd(rng);
        };
        fill(PN_W1, PN_B1 - PN_W1, PN_CIN * PN_KW);
        fill(PN_W2, PN_B2 - PN_W2, PN_CH * PN_KW);
        fill(PN_W3, PN_B3 - PN_W3, PN_CH * PN_KW);
        fill(PN_W4, PN_B4 - PN_W4, PN_CH * PN_KW);
        HIPCHK(hipMemcpy(net->wt.params, w.data(), sizeof(float) * PN_NPARAMS, hipMemcpyHostToDevice));
        pn_pack(net->wt, net->stream);
        HIPCHK(hipStreamSynchronize(net->stream));
        *out = net.release();
    });
}

int lss_proposal_set_params(lss_proposal* net, const float* params, int64_t n) {
    return guarded([&] {
        if (!net || !params || n != PN_NPARAMS) throw LssError(LSS_ERR_ARG, "expected lss_proposal_n_params() values");
        HIPCHK(hipSetDevice(net->dev));
        HIPCHK(hipMemcpy(net->wt.params, params, sizeof(float) * PN_NPARAMS, hipMemcpyHostToDevice));
        HIPCHK(hipMemset(net->d_m, 0, sizeof(float) * PN_NPARAMS));
        HIPCHK(hipMemset(net->d_v, 0, sizeof(float) * PN_NPARAMS));
        net->step = 0;
        pn_pack(net->wt, net->stream);
        HIPCHK(hipStreamSynchronize(net->stream));
    });
}

int lss_proposal_get_params(lss_proposal* net, float* params, int64_t n) {
    return guarded([&] {
        if (!net || !params || n != PN_NPARAMS) throw LssError(LSS_ERR_ARG, "expected lss_proposal_n_params() values");
        HIPCHK(hipSetDevice(net->dev));
        HIPCHK(hipStreamSynchronize(net->stream));
        HIPCHK(hipMemcpy(params, net->wt.params, sizeof(float) * PN_NPARAMS, hipMemcpyDeviceToHost));
    });
}

namespace {
void pn_check_width(const lss_proposal* net, int32_t n) {
    if (n <= 0 || n > net->max_W) throw LssError(LSS_ERR_ARG, "batch width outside 1..max_batch");
}
void pn_upload_xw(lss_proposal* net, const float* xw, int32_t n) {
    HIPCHK(hipMemcpyAsync(net->d_xw, xw, sizeof(float) * (size_t)net->C * n, hipMemcpyHostToDevice, net->stream));
}
void pn_upload_labels(lss_proposal* net, const uint8_t* y, int32_t n) {
    HIPCHK(hipMemsetAsync(net->d_y, 0, (size_t)net->G * net->buf.Wp, net->stream));
    HIPCHK(hipMemcpy2DAsync(net->d_y, net->buf.Wp, y, n, n, net->G, hipMemcpyHostToDevice, net->stream));
}
float pn_loss(lss_proposal* net) {
    const int nb = (net->G * net->buf.Wp + 255) / 256;
    std::vector<float> lp(nb);
    HIPCHK(hipMemcpyAsync(lp.data(), net->buf.lpart, sizeof(float) * nb, hipMemcpyDeviceToHost, net->stream));
    HIPCHK(hipStreamSynchronize(net->stream));
    double l = 0.0;
    for (float v : lp) l += v;
    return (float)l;
}
}  // namespace

int lss_proposal_forward(lss_proposal* net, const float* xw, int32_t n, float* prob) {
    return guarded([&] {
        if (!net || !xw || !prob) throw LssError(LSS_ERR_ARG, "null argument");
        pn_check_width(net, n);
        HIPCHK(hipSetDevice(net->dev));
        pn_upload_xw(net, xw, n);
        pn_infer(net->wt, net->buf, net->d_xw, n, 0, net->C, n, net->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy2DAsync(prob, sizeof(float) * n, net->buf.p, sizeof(float) * net->buf.Wp, sizeof(float) * n,
                                net->G, hipMemcpyDeviceToHost, net->stream));
        HIPCHK(hipStreamSynchronize(net->stream));
    });
}

int lss_proposal_grad(lss_proposal* net, const float* xw, const uint8_t* labels, int32_t n, float pos_weight,
                      float* loss, float* grad) {
    return guarded([&] {
        if (!net || !xw || !labels || !grad) throw LssError(LSS_ERR_ARG, "null argument");
        pn_check_width(net, n);
        HIPCHK(hipSetDevice(net->dev));
        pn_upload_xw(net, xw, n);
        pn_upload_labels(net, labels, n);
        pn_forward(net->wt, net->buf, net->d_xw, n, 0, net->C, n, net->stream);
        pn_backward(net->wt, net->buf, net->d_y, n, pos_weight, net->d_grad, net->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(grad, net->d_grad, sizeof(float) * PN_NPARAMS, hipMemcpyDeviceToHost, net->stream));
        const float l = pn_loss(net);
        if (loss) *loss = l;
    });
}

int lss_proposal_train_step(lss_proposal* net, lss_handle* h, const int16_t* d_samples, int64_t n, int32_t ld,
                            const uint8_t* labels, float pos_weight, float lr, float* loss) {
    return guarded([&] {
        if (!net || !h || !d_samples || !labels) throw LssError(LSS_ERR_ARG, "null argument");
        if (n <= 0 || n > std::min<int64_t>(net->max_W, h->Wmax)) throw LssError(LSS_ERR_ARG, "batch too wide");
        if (h->dev != net->dev) throw LssError(LSS_ERR_ARG, "network and handle on different devices");
        if (pn_groups(h->C) != net->G) throw LssError(LSS_ERR_ARG, "network built for another channel count");
        HIPCHK(hipSetDevice(net->dev));
        set_cmap(h, nullptr, ld);
        BatchPlan bp;
        plan_batch(h, handle_bufs(h), d_samples, nullptr, (int)n, ld, MIN_CW, 0, 1, bp);
        if (bp.has_pre) {
            PreBatch pre{};
            pre.n = 1;
            pre.job[0] = bp.pre;
            launch_preprocess(pre, h->stream);
        }
        WhitenBatch wb{};
        wb.n = 1; wb.start[0] = 0; wb.start[1] = gemm_blocks(bp.wj.mtiles, bp.wj.ntiles); wb.job[0] = bp.wj;
        launch_whiten(wb, nullptr, h->stream);
        HIPCHK(hipStreamSynchronize(h->stream));
        pn_upload_labels(net, labels, (int32_t)n);
        pn_forward(net->wt, net->buf, h->d_Xw, h->XS, XOFF, h->C, (int)n, net->stream);
        pn_backward(net->wt, net->buf, net->d_y, (int)n, pos_weight, net->d_grad, net->stream);
        net->step += 1;
        pn_adam(net->wt, net->d_grad, net->d_m, net->d_v, lr, 0.9f, 0.999f, 1e-8f, net->step, net->stream);
        pn_pack(net->wt, net->stream);
        HIPCHK(hipGetLastError());
        const float l = pn_loss(net);
        if (loss) *loss = l;
    });
}

int lss_proposal_destroy(lss_proposal* net) {
    return guarded([&] { delete net; });
}

int lss_set_proposal(lss_handle* h, lss_proposal* net, float lo, float hi, int32_t reach) {
    return guarded([&] {
        if (!h) throw LssError(LSS_ERR_ARG, "null handle");
        HIPCHK(hipSetDevice(h->dev));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (!net) { h->prop = nullptr; return; }
        if (h->sem != LSS_SEM_KILOSORT) throw LssError(LSS_ERR_ARG, "the proposal gate is defined for Kilosort semantics");
        if (net->dev != h->dev) throw LssError(LSS_ERR_ARG, "network and handle on different devices");
        if (!(lo >= 0.f && lo < hi && hi <= 1.f)) throw LssError(LSS_ERR_ARG, "need 0 <= lo < hi <= 1");
        if (reach < 0 || reach > h->M) throw LssError(LSS_ERR_ARG, "reach outside 0..M");
        const int G = pn_groups(h->C);
        if (G > 64) throw LssError(LSS_ERR_ARG, "at most 1024 channels");
        if (!h->pnb.a0) alloc_pn_buffers(h->pnb, G, h->Wmax);
        const int MT = h->TP / GEMM_BM;
        if (!h->d_gate) {
            const int NBmax = (h->Wmax + 63) / 64;
            h->d_gate = dalloc<uint8_t>((size_t)NBmax * MT);
            // group span of each 64-template tile (internal order, sorted by main channel)
            std::vector<int> glo(MT, 1 << 20), ghi(MT, -(1 << 20));
            for (int tp = 0; tp < h->T; ++tp) {
                const int m = tp / GEMM_BM;
                const int g = h->iU_int[tp] / PN_GROUP;
                glo[m] = std::min(glo[m], g);
                ghi[m] = std::max(ghi[m], g);
            }
            h->d_mt_glo = dalloc<int>(MT); upload(h->d_mt_glo, glo);
            h->d_mt_ghi = dalloc<int>(MT); upload(h->d_mt_ghi, ghi);
            h->d_gate_stats = dalloc<unsigned long long>(4);
        }
        h->gate.mask = h->d_gate;
        h->gate.mt_glo = h->d_mt_glo;
        h->gate.mt_ghi = h->d_mt_ghi;
        h->gate.MT = MT;
        h->gate.lo = lo;
        h->gate.hi = hi;
        h->gate.reach = reach;
        h->gate.stats = h->d_gate_stats;
        h->prop = net;
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free batch entry point.  Its per-batch device buffers (whitened
// batch, PC projection, scores, the peel's exchange words, counters and the
// event staging) are carved from one static device array instead of the
// handle's allocations; the tables stay the handle's.  The head of the array
// holds the tag epoch (its own 256-byte line) and a fixed region for the
// peel's tagged column maxima, so those words are only ever written as tagged
// words, whichever handle used the array before.  Every other buffer is
// written by the batch before it is read, and the per-batch words (counters,
// flags, the epoch) are reset by the whitening launch, so a call needs no
// clearing launch of its own.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[LSS_STATIC_SCRATCH_BYTES];

namespace {

constexpr size_t STATIC_TAG_BYTES = size_t(8) << 20;   // epoch line + tagged column maxima
constexpr int STATIC_MAX_DEVICES = 64;

struct StaticSlot {
    std::once_flag once;
    unsigned char* ptr = nullptr;
    hipError_t rc = hipSuccess;
};
StaticSlot g_static[STATIC_MAX_DEVICES];

// Device address of this device's copy of `scratch` (the host-side name is a
// stand-in, never a device address).  Looked up once per device, with that
// device current.
unsigned char* static_scratch(int dev) {
    if (dev < 0 || dev >= STATIC_MAX_DEVICES) throw LssError(LSS_ERR_ARG, "device ordinal beyond the static table");
    StaticSlot& sl = g_static[dev];
    std::call_once(sl.once, [&] {
        void* p = nullptr;
        sl.rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        sl.ptr = sl.rc == hipSuccess ? static_cast<unsigned char*>(p) : nullptr;
    });
    if (sl.rc != hipSuccess || !sl.ptr)
        throw LssError(LSS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(sl.rc));
    return sl.ptr;
}

// The handle's per-batch buffers laid out in the static array; *need = bytes used.
BatchBufs static_bufs(const lss_handle* h, unsigned char* base, size_t* need) {
    if (h->ptag_words > (int64_t)((STATIC_TAG_BYTES - 256) / sizeof(unsigned long long)))
        throw LssError(LSS_ERR_ARG, "the handle's peel exchange needs " + std::to_string(h->ptag_words * 8) +
                                        " bytes, more than the static array's " +
                                        std::to_string(STATIC_TAG_BYTES - 256));
    return layout_bufs(h, base, STATIC_TAG_BYTES, need);
}

}  // namespace

extern "C" int lss_process_batch_device_static(lss_handle* h, const int16_t* d_samples, int64_t n_samples, int32_t ld,
                                               const int32_t* chan_map, int64_t first_sample, lss_event* ev,
                                               int64_t ev_cap, int64_t* n_ev, float* feat) {
    return guarded([&] {
        if (!h || !n_ev || !d_samples || ev_cap < 0) throw LssError(LSS_ERR_ARG, "null argument");
        *n_ev = 0;
        if (n_samples <= 0 || n_samples > h->Wmax)
            throw LssError(LSS_ERR_ARG, "batch of " + std::to_string(n_samples) + " samples outside 1..max_batch " +
                                            std::to_string(h->Wmax));
        if (h->prop) throw LssError(LSS_ERR_ARG, "the allocation-free entry point does not run a proposal gate");
        if (h->cfg.keep_stages) throw LssError(LSS_ERR_ARG, "the allocation-free entry point keeps no stage copies");
        HIPCHK(hipSetDevice(h->dev));
        const auto t0 = std::chrono::steady_clock::now();
        unsigned char* base = static_scratch(h->dev);
        size_t need = 0;
        const BatchBufs bb = static_bufs(h, base, &need);
        if (need > (size_t)LSS_STATIC_SCRATCH_BYTES)
            throw LssError(LSS_ERR_ARG, "this handle's batch needs " + std::to_string(need) +
                                            " bytes of device scratch; the static array holds " +
                                            std::to_string((size_t)LSS_STATIC_SCRATCH_BYTES) +
                                            " (create the handle with a smaller max_batch or max_events)");
        set_cmap(h, chan_map, ld);
        run_batch(h, bb, d_samples, (int)n_samples, ld, first_sample, false);
        OutHeader hd;
        HIPCHK(hipMemcpy(&hd, bb.hdr, sizeof(hd), hipMemcpyDeviceToHost));
        check_peel(h, hd, (int)n_samples);
        if (hd.count > h->ev_cap)
            throw LssError(LSS_ERR_LIMIT, "the batch has " + std::to_string(hd.count) +
                                              " events, more than the handle's max_events " + std::to_string(h->ev_cap));
        const int64_t n = hd.count, out = std::min<int64_t>(n, ev_cap);
        *n_ev = n;
        if (ev && out > 0) {
            HIPCHK(hipMemcpy(ev, bb.ev, (size_t)out * sizeof(lss_event), hipMemcpyDeviceToHost));
            for (int64_t k = 0; k < out; ++k) ev[k].sample += first_sample;   // the kernel writes column-0-relative samples
        }
        if (feat && out > 0)
            HIPCHK(hipMemcpy(feat, bb.feat, (size_t)out * h->nC * h->K * sizeof(float), hipMemcpyDeviceToHost));
        if (n > ev_cap)
            throw LssError(LSS_ERR_CAPACITY, "event buffer holds " + std::to_string(ev_cap) + " of " +
                                                 std::to_string(n) + " events");
        if (h->pre) h->info.p2p = hd.p2p;
        h->last_pos.clear();
        h->info.events += n;
        h->info.last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
}

namespace lss {
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace lss
