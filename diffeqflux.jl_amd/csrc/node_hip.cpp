// Host driver and C ABI of the NeuralODE hot path (see include/node_hip.h).
//
// The solve loop never waits on the GPU per step: the step controller runs on
// the device (a single-workgroup kernel after each attempt) and every kernel
// reads t/dt/accept state from the device control block, so the host enqueues
// chunks of step attempts and only polls the control block between chunks.
// Kernels of attempts enqueued after the solve has finished return at once.
#include "node_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "node_internal.h"

using namespace nodehip;

namespace {

constexpr int KPAD = 64;    // feature padding (K tile of the BF16 GEMM)
constexpr int BPAD = 128;   // batch padding (largest row tile)
constexpr int SEC = 2 * MAX_PART;  // doubles per partial-sum section

inline int64_t rup(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

const double kB[7] = {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742,
                      -3.290069515436081, 2.324710524099774, 0.0};
const double kBT[7] = {-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995,
                       -0.1447110071732629, 0.5823571654525552, -0.45808210592918697,
                       0.015151515151515152};

struct Layer {
  int in, out, inp, outp, act;
  int64_t w0, b0;
  void* Wr = nullptr;     // [planes][outp][inp]
  void* Wc = nullptr;     // [planes][inp][outp]
  float* bias = nullptr;  // [outp]
  // weight-gradient GEMM plan (rows = inp, cols = outp, K = batch)
  int wg_cfg = 0, wg_rt = 0, wg_ct = 0, wg_ks = 1;
  float* wpart = nullptr; // split-K partials [wg_ks][in*out] when wg_ks > 1
  int64_t wpart_cap = 0;  // floats allocated for wpart
  int wg_ks0 = 1;         // heuristic split-K (autotuning off)
  // column sums of this layer's output gradient per row tile of the data
  // gradient GEMM that produces it ([Bp/32][outp]; hidden layers only)
  float* bpart = nullptr;
  int bpart_rt = 0;
};

struct GemmPlan {
  int cfg, rt, ct;
};

}  // namespace

struct node_handle {
  node_desc desc{};
  int L = 0, D = 0, Dp = 0, B = 0, Bp = 0, prec = 0, esz = 4, planes = 1;
  int wg_prec = 0;  // precision of the weight-gradient products (<= prec)
  bool cnf = false;        // continuous normalizing flow: state [z; log p]
  bool cnf_active = false; // inside node_cnf_forward (probe set)
  bool cnf_adj = false;    // inside node_cnf_backward (tangent chain on)
  bool cnf_fwd = false;    // the last solve was a CNF forward (adjoint available)
  float* probe = nullptr;  // CNF probe e, FP32 [Bp][dm]
  float* fbuf = nullptr;   // CNF model output f, FP32 [Bp][dm]
  std::vector<float> probe_h;  // host copy of the last CNF forward's probe [B][dm]
  // FFJORD variants (input_map flags): regularize adds the state rows r1, r2
  // (r1' = |f|^2, r2' = |e^T df/dz|); the exact trace runs one unit-probe pass
  // per state dimension
  bool cnf_reg = false, cnf_exact = false;
  int cnf_dm = 0;              // z dimension (model output width)
  int npr = 1;                 // probe passes per right-hand side / tangent chains per VJP
  std::vector<void*> LmS;      // exact trace: unit-probe seeds at the last layer (forward)
  std::vector<void*> E0s, AbLs;  // tangent seeds [e_k; 0] and -lambda_dlogp e_k (adjoint)
  float* trc = nullptr;        // exact trace accumulators: trace [Bp], e^T J [Bp][dm]
  float* wacc = nullptr;
  std::vector<float> cnf_lam;  // r1(T), r2(T) of the last forward, [2][B]
  // CNF adjoint (second-order VJP of the Hutchinson term): tangent t_l and
  // its cotangent row-major per layer (operand format), the FP32 second-order
  // sources, and [e; 0] at layer 0's input width
  std::vector<void*> Tn, Ab;
  std::vector<float*> src2;
  void* E0 = nullptr;
  // row stride of the transposed operands (the weight gradients' K): Bp, and
  // 2 Bp on a CNF handle, whose second half holds the tangent chain
  int ldT = 0;
  int64_t P = 0;
  std::vector<Layer> layers;
  hipStream_t stream = nullptr;
  std::string err;
  std::mutex mtx;

  Ctl* ctl = nullptr;
  Ctl* hctl = nullptr;  // pinned mirror
  double* part = nullptr;
  float* tape = nullptr;
  int64_t tape_cap = 0;
  double* fwd_times = nullptr;
  double* save_times = nullptr;
  int64_t save_cap = 0;
  std::vector<double> h_save;  // full list of saved times of the last forward
  float* out_dev = nullptr;
  int64_t out_dev_cap = 0;
  float* scratch = nullptr;
  float* ucur = nullptr;

  void* X0 = nullptr;
  void* X0T = nullptr;
  std::vector<void*> H, HT, S, Dl, DlT;  // per hidden layer (L-1)

  // adjoint
  float* lam = nullptr;
  float* lam_new = nullptr;
  float* kl = nullptr;
  void* Lm = nullptr;
  void* LmT = nullptr;
  double* mu = nullptr;
  float* accb = nullptr;
  float* acce = nullptr;
  float* gf0 = nullptr;
  float* gf1 = nullptr;
  float* dLdout_dev = nullptr;
  int64_t dLdout_cap = 0;
  float* dp32 = nullptr;

  // InterpolatingAdjoint(checkpointing = true): the forward keeps a ring of
  // tape entries plus the checkpoint states (u0, the saved states, u(t1));
  // the adjoint re-solves each interval between consecutive checkpoints into
  // the tape before integrating over it
  bool ckpt = false;          // NODE_OPT_CHECKPOINTING
  bool fwd_ring = false;      // the last forward kept checkpoints only
  int64_t tape_budget = 0;    // NODE_OPT_TAPE_BUDGET: bytes of tape (0: unlimited)
  int rg_vjp_prec = NODE_PREC_FP32;  // NODE_OPT_RAGGED_VJP_PRECISION
  float* ck_u0 = nullptr;     // u(t0), FP32 [B][D]
  float* ck_uT = nullptr;     // u(t1), FP32 [B][D]
  Ctl* ctl_save = nullptr;    // the paused adjoint's control block
  std::vector<double> ck_t;   // checkpoint times t0 < ... < t1
  std::vector<int64_t> ck_src;  // state of each checkpoint: saved index, -1 u0, -2 u(t1)
  int64_t ck_n = 0;           // steps of the interval on the tape
  double ck_last_dt = 0.0;    // last step of the previous re-solve (next one's initial dt)
  int64_t ck_nf = 0;          // RHS evaluations of the re-solves of the last backward

  void* na_scratch = nullptr;  // this device's no-alloc array (node_rhs_device_noalloc)

  bool have_params = false, have_forward = false;
  node_solve_opts fwd_opts{};
  node_stats fwd_stats{};
  int64_t n_fwd = 0;
  int64_t nsaved = 0;
  bool everystep = false;

  std::vector<void*> allocs;

  // autotuned tile configuration per (role, layer); -1 = heuristic
  int tcfg[6][MAXL];
  int tks[6][MAXL];       // split-K of the stage epilogues (serial fixup)
  bool tuned = false, autotune = true;
  bool tuned_data = false;  // plans of every role but the weight gradients are tuned
  int sense = NODE_SENSE_INTERPOLATING;  // adjoint of node_backward (NODE_SENSE_*)
  // BacksolveAdjoint state: u at the step start / candidate, 8 stage slots of
  // f(u), f(u_s) of the current stage
  float* bs_u = nullptr;
  float* bs_unew = nullptr;
  float* bs_ku = nullptr;
  float* bs_fu = nullptr;
  int tune_role = -1, tune_layer = -1, tune_cfg = -1, tune_ks = 1;  // override while timing
  float* tune_wpart = nullptr;
  float* fix_ws = nullptr;  // split-K fixup workspace / per-tile counters
  int* fix_cnt = nullptr;

  // captured step-attempt graphs (re-captured when a captured pointer changes)
  bool use_graphs = true;
  hipGraphExec_t fwd_graph = nullptr, adj_graph = nullptr;
  const void* fwd_graph_sig[3] = {nullptr, nullptr, nullptr};  // tape, save times, outputs
  // tape, save times, cotangents, saved states (BacksolveAdjoint checkpoints)
  const void* adj_graph_sig[4] = {nullptr, nullptr, nullptr, nullptr};
  int graph_attempts = 4;

  // ragged batches: flat parameters and the last per-trajectory forward
  float* pflat = nullptr;
  struct Ragged {
    int64_t n = 0, n_cap = 0, ns = 0, ns_cap = 0;
    double* t1 = nullptr;
    float* u0 = nullptr;
    int64_t* soff = nullptr;
    double* st = nullptr;
    float* out = nullptr;
    int* ret = nullptr;
    int* tstats = nullptr;
    int* last = nullptr;
    int* ctr = nullptr;                     // [0] queue, [1] tape records
    unsigned long long* stats = nullptr;    // nf, naccept, nreject
    double* rec_t = nullptr;
    int* rec_prev = nullptr;
    float* rec_data = nullptr;
    int64_t rec_cap = 0;
    bool have = false;
    node_solve_opts opts{};
    std::vector<int> h_ret;
  } rg;
  int n_cu = 0;
  struct RaggedAdj {
    float* dLdout = nullptr; int64_t dLdout_cap = 0;
    float* dLdu0 = nullptr; int64_t dLdu0_cap = 0;
    float* rows = nullptr; int64_t rows_cap = 0;
    float* tmp = nullptr; int64_t tmp_cap = 0;
    float* sbuf = nullptr; int64_t sbuf_cap = 0;
    double* dp = nullptr; int64_t dp_cap = 0;
    int* order = nullptr; int64_t order_cap = 0;
    double* part = nullptr; int64_t part_cap = 0;
  } ra;

  // profiling (node_prof_enable)
  uint64_t prof_mask = 0;
  struct ProfRec { int key; hipEvent_t a, b; };
  std::vector<ProfRec> prof;
  std::vector<hipEvent_t> ev_pool;
};

namespace {

int fail(node_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(h, NODE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

int dalloc(node_handle* h, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess)
    return fail(h, NODE_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemsetAsync(*p, 0, bytes, h->stream);
  if (e != hipSuccess)
    return fail(h, NODE_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
  h->allocs.push_back(*p);
  return 0;
}
template <typename P>
int dalloc_t(node_handle* h, P** p, size_t bytes) {
  void* v = nullptr;
  int rc = dalloc(h, &v, bytes);
  *p = (P*)v;
  return rc;
}

void dfree(node_handle* h, void* p) {
  if (!p) return;
  auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
  if (it != h->allocs.end()) h->allocs.erase(it);
  (void)hipFree(p);
}

// Tile configuration with >= 240 workgroups if one exists (largest tile
// first), else the one with the most workgroups.  min_cfg > 0 skips the
// 128x128 tiles (the stage epilogues use at most 64 columns per tile).
bool cfg_allowed(int c, int min_cfg) { return c >= min_cfg && (min_cfg == 0 || gemm_cfg_tile_cols(c) <= 64); }
GemmPlan plan(int rows, int cols, int min_cfg = 0) {
  int best = -1, best_tiles = -1;
  for (int c = min_cfg; c < gemm_cfg_count(); ++c) {
    if (!cfg_allowed(c, min_cfg)) continue;
    const int br = gemm_cfg_tile_rows(c), bc = gemm_cfg_tile_cols(c);
    if (rows % br || cols % bc) continue;
    const int t = (rows / br) * (cols / bc);
    if (t >= 240) return {c, rows / br, cols / bc};
    if (t > best_tiles) { best = c; best_tiles = t; }
  }
  return {best, rows / gemm_cfg_tile_rows(best), cols / gemm_cfg_tile_cols(best)};
}

enum ProfRole { PR_FWD_HIDDEN = 0, PR_FWD_STAGE = 1, PR_ADJ_HIDDEN = 2, PR_BWD_DATA = 3,
                PR_WGRAD = 4, PR_BWD_LAST = 5 };

hipEvent_t prof_event(node_handle* h) {
  if (!h->ev_pool.empty()) {
    hipEvent_t e = h->ev_pool.back();
    h->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// RAII pair of events around one launch when its key is selected
struct ProfScope {
  node_handle* h;
  int key;
  hipEvent_t a = nullptr;
  ProfScope(node_handle* hh, int role, int layer) : h(hh), key(role * 8 + layer) {
    if (key < 64 && ((h->prof_mask >> key) & 1ull)) {
      a = prof_event(h);
      (void)hipEventRecord(a, h->stream);
    }
  }
  ~ProfScope() {
    if (!a) return;
    hipEvent_t b = prof_event(h);
    (void)hipEventRecord(b, h->stream);
    h->prof.push_back({key, a, b});
  }
};

StateGeom geom(const node_handle* h) {
  StateGeom g;
  g.B = h->B; g.Bp = h->Bp; g.D = h->D; g.Dp = h->Dp;
  g.map = h->desc.input_map; g.prec = h->prec;
  g.sv = (int64_t)h->Bp * h->D;
  g.ent = 8 * g.sv;
  g.dm = h->cnf ? h->cnf_dm : h->D;
  g.tcat = h->desc.input_map == NODE_MAP_CNF_TIME ? 1 : 0;
  g.x0w = h->layers[0].inp;
  return g;
}

int validate_desc(const node_desc* d, std::string& msg) {
  if (!d) { msg = "desc is NULL"; return NODE_ERR_INVALID; }
  if (d->n_layers < 1 || d->n_layers > NODE_HIP_MAX_LAYERS) { msg = "n_layers must be 1..8"; return NODE_ERR_INVALID; }
  for (int i = 0; i <= d->n_layers; ++i)
    if (d->widths[i] < 1 || d->widths[i] > 65536) { msg = "widths must be 1..65536"; return NODE_ERR_INVALID; }
  for (int i = 0; i < d->n_layers; ++i)
    if (d->acts[i] < 0 || d->acts[i] > 4) { msg = "unknown activation"; return NODE_ERR_INVALID; }
  const int base_map = d->input_map & ~(NODE_MAP_CNF_REGULARIZE | NODE_MAP_CNF_EXACT_TRACE);
  if (base_map == NODE_MAP_CNF || base_map == NODE_MAP_CNF_TIME) {
    const int dm = d->widths[d->n_layers];
    const int extra = (d->input_map & NODE_MAP_CNF_REGULARIZE) ? 3 : 1;
    if (d->widths[0] != dm + (base_map == NODE_MAP_CNF_TIME ? 1 : 0)) { msg = "CNF model: input width must be D (or D + 1 with time)"; return NODE_ERR_INVALID; }
    if (d->n_layers < 2) { msg = "CNF model needs at least 2 Dense layers"; return NODE_ERR_INVALID; }
    if (d->acts[d->n_layers - 1] != NODE_ACT_IDENTITY) { msg = "CNF model: the last layer must be linear"; return NODE_ERR_INVALID; }
    if (dm + extra > 64) { msg = "CNF state dimension (D + 1, D + 3 with regularize) must be <= 64"; return NODE_ERR_INVALID; }
  } else if (base_map != d->input_map) {
    msg = "NODE_MAP_CNF_REGULARIZE / NODE_MAP_CNF_EXACT_TRACE apply to CNF maps only"; return NODE_ERR_INVALID;
  } else {
    if (d->widths[0] != d->widths[d->n_layers]) { msg = "output width must equal input width (ODE state dimension)"; return NODE_ERR_INVALID; }
    if (d->input_map != NODE_MAP_IDENTITY && d->input_map != NODE_MAP_CUBE) { msg = "unknown input_map"; return NODE_ERR_INVALID; }
  }
  if (d->precision < NODE_PREC_FP32 || d->precision > NODE_PREC_BF16X3) { msg = "unknown precision"; return NODE_ERR_INVALID; }
  if (d->batch < 1 || d->batch > (int64_t)1 << 24) { msg = "batch must be 1..2^24"; return NODE_ERR_INVALID; }
  return 0;
}

// ---------------------------------------------------------------- RHS / VJP
// Tile plan of one GEMM role: the autotuned choice when there is one, else the
// occupancy heuristic.  min_cfg 1 for the stage epilogues (<= 64 columns).
GemmPlan plan_role(node_handle* h, int role, int layer, int rows, int cols, int min_cfg) {
  if (h->tune_role == role && h->tune_layer == layer && h->tune_cfg >= 0) {
    const int c = h->tune_cfg;
    return {c, rows / gemm_cfg_tile_rows(c), cols / gemm_cfg_tile_cols(c)};
  }
  const int c = h->tcfg[role][layer];
  if (c >= 0) return {c, rows / gemm_cfg_tile_rows(c), cols / gemm_cfg_tile_cols(c)};
  return plan(rows, cols, min_cfg);
}

// The skinny stage GEMMs (last forward layer, layer-1 data gradient) run
// their per-element stage work in a separate elementwise kernel over split-K
// partials when the fused epilogue would have fewer than 2048 tiles of 32x64
// (its ~14 dependent loads per element are then memory-latency bound); larger
// batches keep the fused epilogue, which saves the partials' round trip.
bool stage_split(const node_handle* h, int role) {
  if (h->cnf && role == PR_FWD_STAGE) return false;
  if (role == PR_BWD_LAST && (h->sense != NODE_SENSE_INTERPOLATING || h->cnf))
    return true;  // backsolve / CNF adjoint stage: elementwise kernel only
  const int cols = role == PR_BWD_LAST ? h->Dp : h->layers[h->L - 1].outp;
  return (int64_t)(h->Bp / 32) * (cols / 64) < 2048;
}

// split-K factor of a stage-epilogue role (1 without a fixup workspace)
int stage_ks(node_handle* h, int role, int layer) {
  if (!h->fix_ws) return 1;
  if (h->tune_role == role && h->tune_layer == layer && h->tune_cfg >= 0) return h->tune_ks;
  return h->tks[role][layer];
}

// act' of identity, softplus and relu is a function of the layer output: the
// adjoint recompute then stores no act' and the data gradient derives it
bool act_from_output(int act) {
  return act == NODE_ACT_IDENTITY || act == NODE_ACT_SOFTPLUS || act == NODE_ACT_RELU;
}

// hidden layer l: forward (Y) or adjoint recompute (Y, Y^T, act')
void launch_hidden(node_handle* h, int l, int gate, bool adjoint) {
  const Layer& ly = h->layers[l];
  GemmArgs g{};
  g.A = l == 0 ? h->X0 : h->H[l - 1];
  g.lda = ly.inp; g.B = ly.Wr; g.ldb = ly.inp; g.K = ly.inp;
  g.pa = (int64_t)h->Bp * ly.inp; g.pb = (int64_t)ly.outp * ly.inp; g.ksplit = 1;
  const int role = adjoint ? PR_ADJ_HIDDEN : PR_FWD_HIDDEN;
  GemmPlan p = plan_role(h, role, l, h->Bp, ly.outp, 0);
  g.row_tiles = p.rt; g.col_tiles = p.ct;
  const int64_t ph = (int64_t)h->Bp * ly.outp;
  EpiHidden e{ly.bias, ly.act, ly.out, ly.outp, h->ldT, h->H[l], adjoint ? h->HT[l] : nullptr,
              adjoint && !act_from_output(ly.act) ? h->S[l] : nullptr, ph,
              (int64_t)h->ldT * ly.outp};
  ProfScope ps(h, role, l);
  launch_gemm_hidden(h->prec, p.cfg, g, e, h->ctl, gate, h->stream);
}

// last layer with the Tsit5 stage epilogue
void launch_fwd_stage(node_handle* h, int stage, int mode, int gate, double* part,
                      float* plain_out = nullptr) {
  const int l = h->L - 1;
  const Layer& ly = h->layers[l];
  GemmArgs g{};
  g.A = l == 0 ? h->X0 : h->H[l - 1];
  g.lda = ly.inp; g.B = ly.Wr; g.ldb = ly.inp; g.K = ly.inp;
  g.pa = (int64_t)h->Bp * ly.inp; g.pb = (int64_t)ly.outp * ly.inp; g.ksplit = 1;
  GemmPlan p = plan_role(h, PR_FWD_STAGE, l, h->Bp, ly.outp, 1);
  g.row_tiles = p.rt; g.col_tiles = p.ct;
  g.ksplit = stage_ks(h, PR_FWD_STAGE, l);
  const bool split = stage_split(h, PR_FWD_STAGE);
  g.ws = split ? nullptr : h->fix_ws;
  g.cnt = split ? nullptr : h->fix_cnt;
  EpiFwdStage e{};
  e.bias = ly.bias; e.act = ly.act; e.D = h->D; e.B = h->B; e.stage = stage; e.mode = mode;
  e.tape = h->tape; e.ent = 8 * (int64_t)h->Bp * h->D; e.sv = (int64_t)h->Bp * h->D;
  e.X0 = h->X0; e.ldx = h->Dp; e.map = h->desc.input_map;
  e.scratch = plain_out ? plain_out : h->scratch; e.part = part;
  e.px = (int64_t)h->Bp * h->Dp;
  ProfScope ps(h, PR_FWD_STAGE, l);
  if (split) {
    const EpiPartial ep{h->fix_ws, ly.outp, (int64_t)h->Bp * ly.outp};
    launch_gemm_partial(h->prec, p.cfg, g, ep, h->ctl, gate, h->stream);
    launch_fwd_stage_kernel(h->prec, e, h->fix_ws, g.ksplit, ep.pstride, ly.outp, h->Bp, ly.outp,
                            h->ctl, gate, h->stream);
  } else {
    launch_gemm_fwdstage(h->prec, p.cfg, g, e, h->ctl, gate, h->stream);
  }
}

// One batched right-hand side: X0 -> layers -> last-layer epilogue.
void cnf_eval(node_handle* h, int stage, int mode, int gate, double* part);
int cnf_last_tiles(node_handle* h);

void rhs_eval(node_handle* h, int stage, int mode, int gate, double* part) {
  if (h->cnf) { cnf_eval(h, stage, mode, gate, part); return; }
  for (int l = 0; l < h->L - 1; ++l) launch_hidden(h, l, gate, false);
  launch_fwd_stage(h, stage, mode, gate, part);
}

int rhs_last_tiles(node_handle* h) {
  if (h->cnf) return cnf_last_tiles(h);
  const Layer& ly = h->layers[h->L - 1];
  if (stage_split(h, PR_FWD_STAGE)) return stage_grid(h->Bp, ly.outp);
  GemmPlan p = plan_role(h, PR_FWD_STAGE, h->L - 1, h->Bp, ly.outp, 1);
  return p.rt * p.ct;
}

// data gradient of layer l >= 1: delta_{l-1} = (delta_l W_l) * act'
void launch_bwd_data(node_handle* h, int l, int gate, bool bias_sums = true) {
  const Layer& ly = h->layers[l];
  GemmArgs g{};
  g.A = l == h->L - 1 ? h->Lm : h->Dl[l];
  g.lda = ly.outp; g.B = ly.Wc; g.ldb = ly.outp; g.K = ly.outp;
  g.pa = (int64_t)h->Bp * ly.outp; g.pb = (int64_t)ly.inp * ly.outp; g.ksplit = 1;
  GemmPlan p = plan_role(h, PR_BWD_DATA, l, h->Bp, ly.inp, 0);
  g.row_tiles = p.rt; g.col_tiles = p.ct;
  const int64_t pd = (int64_t)h->Bp * ly.inp;
  Layer& lo = h->layers[l - 1];
  float* bsum = bias_sums ? lo.bpart : nullptr;
  if (bsum) lo.bpart_rt = p.rt;
  EpiBwdData e{ly.in, ly.inp, h->ldT, h->S[l - 1], h->Dl[l - 1], h->DlT[l - 1], pd,
               (int64_t)h->ldT * ly.inp, bsum, h->H[l - 1], pd,
               act_from_output(lo.act) ? lo.act : -1};
  ProfScope ps(h, PR_BWD_DATA, l);
  if (h->cnf_adj) {  // plus the Hutchinson term's second-order source
    const EpiBwdSrc es{e, h->src2[l - 1]};
    launch_gemm_bwdsrc(h->prec, p.cfg, g, es, h->ctl, gate, h->stream);
  } else {
    launch_gemm_bwddata(h->prec, p.cfg, g, e, h->ctl, gate, h->stream);
  }
}

// pointer to tangent slice k of a CNF handle's transposed operand (columns
// (1 + k) Bp ..: the primal rows come first, then one slice per probe chain)
void* tan_half(const node_handle* h, void* T, int k = 0) {
  return (char*)T + (size_t)(1 + k) * h->Bp * h->esz;
}

// CNF adjoint, tangent forward of hidden layer l: t_l = act'_l * (t_{l-1} W_l^T)
// with t_{-1} = [e; 0], row-major and transposed (second half of H_l^T)
void launch_tan_fwd(node_handle* h, int l, int gate, int k = 0) {
  const Layer& ly = h->layers[l];
  GemmArgs g{};
  g.A = l == 0 ? h->E0s[k] : h->Tn[l - 1];
  g.lda = ly.inp; g.B = ly.Wr; g.ldb = ly.inp; g.K = ly.inp;
  g.pa = (int64_t)h->Bp * ly.inp; g.pb = (int64_t)ly.outp * ly.inp; g.ksplit = 1;
  GemmPlan p = plan_role(h, PR_ADJ_HIDDEN, l, h->Bp, ly.outp, 0);
  g.row_tiles = p.rt; g.col_tiles = p.ct;
  const int64_t ph = (int64_t)h->Bp * ly.outp;
  EpiBwdData e{ly.out, ly.outp, h->ldT, h->S[l], h->Tn[l], tan_half(h, h->HT[l], k), ph,
               (int64_t)h->ldT * ly.outp, nullptr, h->H[l], ph,
               act_from_output(ly.act) ? ly.act : -1};
  ProfScope ps(h, PR_ADJ_HIDDEN, l);
  launch_gemm_bwddata(h->prec, p.cfg, g, e, h->ctl, gate, h->stream);
}

// CNF adjoint, tangent backward through layer l >= 1 into layer l - 1:
// abar_{l-1} = (abar_l W_l) act'_{l-1} (+ its second-order source)
void launch_tan_bwd(node_handle* h, int l, int gate, int k = 0) {
  const Layer& ly = h->layers[l];
  const Layer& lo = h->layers[l - 1];
  GemmArgs g{};
  g.A = l == h->L - 1 ? h->AbLs[k] : h->Ab[l];
  g.lda = ly.outp; g.B = ly.Wc; g.ldb = ly.outp; g.K = ly.outp;
  g.pa = (int64_t)h->Bp * ly.outp; g.pb = (int64_t)ly.inp * ly.outp; g.ksplit = 1;
  GemmPlan p = plan_role(h, PR_BWD_DATA, l, h->Bp, ly.inp, 0);
  g.row_tiles = p.rt; g.col_tiles = p.ct;
  const int64_t pd = (int64_t)h->Bp * ly.inp;
  EpiTanBwd e{};
  e.C = ly.in; e.ldo = ly.inp; e.ldt = h->ldT; e.act = lo.act;
  e.Y = h->H[l - 1]; e.S = h->S[l - 1]; e.Tf = h->Tn[l - 1]; e.py = e.ptf = pd;
  // (abar_0 row-major feeds no further tangent GEMM: only its transposed copy)
  e.Ab = l - 1 >= 1 ? h->Ab[l - 1] : nullptr;
  e.AbT = tan_half(h, h->DlT[l - 1], k);
  e.pab = pd; e.pabt = (int64_t)h->ldT * ly.inp;
  e.src = h->src2[l - 1];
  e.src_acc = k > 0;  // the second-order sources of all probe chains add up
  ProfScope ps(h, PR_BWD_DATA, l);
  launch_gemm_tanbwd(h->prec, p.cfg, g, e, h->ctl, gate, h->stream);
}

// weight gradient of layer l (K = batch), folded into the mu accumulators or
// written as split-K partials
void launch_wgrad(node_handle* h, int l, int gate, const EpiWgrad& wt) {
  const Layer& ly = h->layers[l];
  GemmArgs g{};
  // (K = ldT: on a CNF handle the second half adds the tangent chain's t^T abar)
  g.A = l == 0 ? h->X0T : h->HT[l - 1];
  g.lda = h->ldT;
  g.B = l == h->L - 1 ? h->LmT : h->DlT[l];
  g.ldb = h->ldT;
  g.K = h->ldT;
  g.pa = (int64_t)ly.inp * h->ldT;
  g.pb = (int64_t)ly.outp * h->ldT;
  int cfg = ly.wg_cfg, ks = ly.wg_ks;
  float* wpart = ly.wpart;
  if (h->tune_role == PR_WGRAD && h->tune_layer == l && h->tune_cfg >= 0) {
    cfg = h->tune_cfg;
    ks = h->tune_ks;
    wpart = h->tune_wpart;
  }
  g.row_tiles = ly.inp / gemm_cfg_tile_rows(cfg);
  g.col_tiles = ly.outp / gemm_cfg_tile_cols(cfg);
  g.ksplit = ks;
  EpiWgrad e = wt;
  e.in = ly.in; e.out = ly.out; e.w0 = ly.w0;
  e.wpart = ks > 1 ? wpart : nullptr;
  ProfScope ps(h, PR_WGRAD, l);
  launch_gemm_wgrad(h->wg_prec, cfg, g, e, h->ctl, gate, h->stream);
}

// layer-1 data gradient with the adjoint stage epilogue
void launch_bwd_last(node_handle* h, int stage, int mode, int gate, double* part,
                     float* plain_out) {
  const int L = h->L;
  const Layer& l0 = h->layers[0];
  GemmArgs g{};
  g.A = L == 1 ? h->Lm : h->Dl[0];
  g.lda = l0.outp; g.B = l0.Wc; g.ldb = l0.outp; g.K = l0.outp;
  g.pa = (int64_t)h->Bp * l0.outp; g.pb = (int64_t)l0.inp * l0.outp; g.ksplit = 1;
  GemmPlan p = plan_role(h, PR_BWD_LAST, 0, h->Bp, h->Dp, 2);
  g.row_tiles = p.rt; g.col_tiles = p.ct;
  g.ksplit = stage_ks(h, PR_BWD_LAST, 0);
  const bool split = stage_split(h, PR_BWD_LAST);
  g.ws = split ? nullptr : h->fix_ws;
  g.cnt = split ? nullptr : h->fix_cnt;
  EpiBwdLast e{};
  e.D = h->D; e.B = h->B; e.stage = stage; e.mode = mode;
  e.kl = h->kl; e.sv = (int64_t)h->Bp * h->D;
  e.lam = h->lam; e.lam_new = h->lam_new; e.Ucur = h->ucur;
  e.L = h->Lm; e.LT = h->LmT; e.ldl = h->Dp; e.ldlt = h->ldT;
  e.X0 = h->X0; e.X0T = h->X0T; e.ldx = h->Dp; e.ldxt = h->ldT; e.map = h->desc.input_map;
  e.tape = h->tape; e.ent = 8 * (int64_t)h->Bp * h->D;
  e.part = part; e.plain_out = plain_out;
  e.pl = e.px = (int64_t)h->Bp * h->Dp;
  e.plt = e.pxt = (int64_t)h->ldT * h->Dp;
  if (h->cnf) {
    const StateGeom sg = geom(h);
    e.cnf = 1; e.dm = sg.dm; e.tcat = sg.tcat;
  }
  if (h->sense != NODE_SENSE_INTERPOLATING) {
    e.fu = h->bs_fu; e.ku = h->bs_ku; e.ubs = h->bs_u; e.ubs_new = h->bs_unew;
  }
  ProfScope ps(h, PR_BWD_LAST, 0);
  if (split) {
    const EpiPartial ep{h->fix_ws, h->Dp, (int64_t)h->Bp * h->Dp};
    launch_gemm_partial(h->prec, p.cfg, g, ep, h->ctl, gate, h->stream);
    launch_bwd_stage_kernel(h->prec, e, h->fix_ws, g.ksplit, ep.pstride, h->Dp, h->Bp, h->Dp,
                            h->ctl, gate, h->stream);
  } else {
    launch_gemm_bwdlast(h->prec, p.cfg, g, e, h->ctl, gate, h->stream);
  }
}

int vjp_last_tiles(node_handle* h) {
  if (stage_split(h, PR_BWD_LAST)) return stage_grid(h->Bp, h->Dp);
  GemmPlan p = plan_role(h, PR_BWD_LAST, 0, h->Bp, h->Dp, 2);
  return p.rt * p.ct;
}

void launch_fold_all(node_handle* h, int gate, const EpiWgrad& wt) {
  FoldArgs fa{};
  fa.nl = h->L;
  for (int l = 0; l < h->L; ++l) {
    const Layer& ly = h->layers[l];
    fa.dT[l] = l == h->L - 1 ? h->LmT : h->DlT[l];
    fa.dplane[l] = (int64_t)ly.outp * h->ldT;
    fa.outs[l] = ly.out;
    fa.b0s[l] = ly.b0;
    fa.bias_rows += ly.out;
    fa.bpart[l] = ly.bpart_rt > 0 ? ly.bpart : nullptr;
    fa.bpart_rt[l] = ly.bpart_rt;
    fa.bpart_ld[l] = ly.outp;
    if (ly.wg_ks > 1) {
      fa.sk_w0[fa.nsk] = ly.w0;
      fa.sk_n[fa.nsk] = (int64_t)ly.in * ly.out;
      fa.sk_part[fa.nsk] = ly.wpart;
      fa.sk_split[fa.nsk] = ly.wg_ks;
      fa.nsk++;
    }
  }
  fa.ldt = h->ldT;
  fa.Bp = h->Bp;
  launch_fold(h->prec, fa, wt, h->ctl, gate, h->stream);
}

// One adjoint stage: forward recompute at the interpolated u, data and weight
// gradients, bias gradients, and the layer-1 epilogue (k^lambda + next stage).
void vjp_eval(node_handle* h, int stage, int mode, int gate, int wflags, float cb, float ce,
              float cb1, float ce1, double* part, float* plain_out, float* gptr) {
  const int L = h->L;
  for (int l = 0; l < L - 1; ++l) launch_hidden(h, l, gate, true);
  if (h->cnf_adj) {  // tangent chains of the trace term (one per probe), forward then reverse
    for (int k = 0; k < h->npr; ++k) {
      for (int l = 0; l < L - 1; ++l) launch_tan_fwd(h, l, gate, k);
      for (int l = L - 1; l >= 1; --l) launch_tan_bwd(h, l, gate, k);
    }
  }
  // BacksolveAdjoint: f(u_s) of this stage (the last layer on the recomputed
  // hidden state) for u's own Runge-Kutta stages
  if (h->sense != NODE_SENSE_INTERPOLATING && mode != SM_PLAIN)
    launch_fwd_stage(h, 0, SM_PLAIN, gate, nullptr, h->bs_fu);
  EpiWgrad wt{};
  wt.accb = h->accb; wt.acce = h->acce; wt.gf0 = h->gf0; wt.gf1 = h->gf1; wt.gptr = gptr;
  wt.cb = cb; wt.ce = ce; wt.cb1 = cb1; wt.ce1 = ce1; wt.flags = wflags;
  for (int l = L - 1; l >= 1; --l) {
    launch_bwd_data(h, l, gate);
    launch_wgrad(h, l, gate, wt);
  }
  launch_wgrad(h, 0, gate, wt);
  launch_fold_all(h, gate, wt);
  launch_bwd_last(h, stage, mode, gate, part, plain_out);
}

int ensure_tape(node_handle* h, int64_t entries) {
  if (entries <= h->tape_cap) return 0;
  int64_t cap = std::max<int64_t>(entries, std::max<int64_t>(16, 2 * h->tape_cap));
  const int64_t ent_bytes = 8 * (int64_t)h->Bp * h->D * sizeof(float);
  if (h->tape_budget > 0) {
    const int64_t lim = h->tape_budget / ent_bytes;
    if (entries > lim)
      return fail(h, NODE_ERR_CAPACITY,
                  "the forward tape needs " + std::to_string(entries) + " entries of " +
                      std::to_string(ent_bytes) + " bytes, beyond the tape budget (" +
                      std::to_string(h->tape_budget) +
                      " bytes); InterpolatingAdjoint(checkpointing = true) keeps only the "
                      "checkpoints and re-solves one interval at a time");
    cap = std::min(cap, lim);
  }
  float* nt = nullptr;
  double* nft = nullptr;
  if (int rc = dalloc_t(h, &nt, (size_t)(cap * ent_bytes))) return rc;
  if (int rc = dalloc_t(h, &nft, (size_t)((cap + 1) * sizeof(double)))) return rc;
  if (h->tape) {
    HIPCHK(h, hipMemcpyAsync(nt, h->tape, h->tape_cap * ent_bytes, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(nft, h->fwd_times, (h->tape_cap + 1) * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    dfree(h, h->tape);
    dfree(h, h->fwd_times);
  }
  h->tape = nt;
  h->fwd_times = nft;
  h->tape_cap = cap;
  return 0;
}

int ensure_wpart(node_handle* h) {
  for (Layer& ly : h->layers) {
    const int64_t need = ly.wg_ks > 1 ? (int64_t)ly.wg_ks * ly.in * ly.out : 0;
    if (need > ly.wpart_cap) {
      dfree(h, ly.wpart);
      ly.wpart = nullptr;
      if (int rc = dalloc_t(h, &ly.wpart, (size_t)need * sizeof(float))) return rc;
      ly.wpart_cap = need;
    }
  }
  return 0;
}

int ensure_adjoint(node_handle* h);

// Average duration (ms) of `reps` back-to-back launches after one warm-up.
template <typename F>
float time_launches(node_handle* h, F&& f, int reps) {
  hipEvent_t a = prof_event(h), b = prof_event(h);
  f();
  (void)hipEventRecord(a, h->stream);
  for (int r = 0; r < reps; ++r) f();
  (void)hipEventRecord(b, h->stream);
  (void)hipEventSynchronize(b);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  h->ev_pool.push_back(a);
  h->ev_pool.push_back(b);
  return ms / reps;
}

// Autotuner: times every valid tile configuration (and split-K factor for the
// weight gradients) of every GEMM role on this handle's own buffers and shapes,
// and keeps the fastest.  Runs once per handle, before its first solve.
int autotune_impl(node_handle* h) {
  if (int rc = ensure_adjoint(h)) return rc;
  const int L = h->L;
  const int reps = 5;
  int64_t maxio = 0;
  for (const Layer& ly : h->layers) maxio = std::max<int64_t>(maxio, (int64_t)ly.in * ly.out);
  float* tmp = nullptr;
  if (int rc = dalloc_t(h, &tmp, (size_t)16 * maxio * sizeof(float))) return rc;  // split-K <= 16
  const uint64_t saved_mask = h->prof_mask;
  h->prof_mask = 0;
  const bool data = !h->tuned_data;
  auto tune_slot = [&](int role, int layer, int rows, int cols, int min_cfg, auto&& launch) {
    if (!data) return;
    float best = 1e30f;
    int bc = -1;
    for (int c = min_cfg; c < gemm_cfg_count(); ++c) {
      if (rows % gemm_cfg_tile_rows(c) || cols % gemm_cfg_tile_cols(c)) continue;
      h->tune_role = role; h->tune_layer = layer; h->tune_cfg = c; h->tune_ks = 1;
      const float t = time_launches(h, launch, reps);
      if (t < best) { best = t; bc = c; }
    }
    h->tune_role = h->tune_layer = h->tune_cfg = -1;
    h->tcfg[role][layer] = bc;
  };
  for (int l = 0; l < L - 1; ++l) {
    const Layer& ly = h->layers[l];
    tune_slot(PR_FWD_HIDDEN, l, h->Bp, ly.outp, 0, [&] { launch_hidden(h, l, GATE_ALWAYS, false); });
    tune_slot(PR_ADJ_HIDDEN, l, h->Bp, ly.outp, 0, [&] { launch_hidden(h, l, GATE_ALWAYS, true); });
  }
  // (the adjoint stage epilogue keeps five value tiles live: its 128-row
  // form spills, so it starts at the 64-row tiles)
  auto tune_stage = [&](int role, int layer, int cols, int K, auto&& launch) {
    if (!data) return;
    float best = 1e30f;
    int bc = -1, bks = 1;
    const int kt = (int)((int64_t)K * h->esz / 128);
    for (int c = role == PR_BWD_LAST ? 2 : 1; c < gemm_cfg_count(); ++c) {
      if (!cfg_allowed(c, 1)) continue;
      if (h->Bp % gemm_cfg_tile_rows(c) || cols % gemm_cfg_tile_cols(c)) continue;
      for (int ks = 1; ks <= 8; ks *= 2) {
        if (kt % ks) continue;
        h->tune_role = role; h->tune_layer = layer; h->tune_cfg = c; h->tune_ks = ks;
        const float t = time_launches(h, launch, reps);
        if (t < best) { best = t; bc = c; bks = ks; }
      }
    }
    h->tune_role = h->tune_layer = h->tune_cfg = -1;
    h->tcfg[role][layer] = bc;
    h->tks[role][layer] = bks;
  };
  tune_stage(PR_FWD_STAGE, L - 1, h->layers[L - 1].outp, h->layers[L - 1].inp,
             [&] { launch_fwd_stage(h, 0, SM_PLAIN, GATE_ALWAYS, nullptr); });
  for (int l = 1; l < L; ++l)
    tune_slot(PR_BWD_DATA, l, h->Bp, h->layers[l].inp, 0, [&] { launch_bwd_data(h, l, GATE_ALWAYS); });
  tune_stage(PR_BWD_LAST, 0, h->Dp, h->layers[0].outp,
             [&] { launch_bwd_last(h, 0, SM_PLAIN, GATE_ALWAYS, nullptr, h->scratch); });
  // weight gradients: tile x split-K; a split pays for its partial sums in the fold
  EpiWgrad wt{};
  wt.accb = h->accb; wt.acce = h->acce; wt.gf0 = h->gf0; wt.gf1 = h->gf1; wt.gptr = h->accb;
  wt.flags = WG_STORE_PTR;
  const int ktiles = (int)((int64_t)h->Bp * h->esz / 128);
  for (int l = 0; l < L; ++l) {
    Layer& ly = h->layers[l];
    float best = 1e30f;
    int bc = ly.wg_cfg, bks = ly.wg_ks;
    for (int c = 0; c < gemm_cfg_count(); ++c) {
      if (ly.inp % gemm_cfg_tile_rows(c) || ly.outp % gemm_cfg_tile_cols(c)) continue;
      const int tiles = (ly.inp / gemm_cfg_tile_rows(c)) * (ly.outp / gemm_cfg_tile_cols(c));
      for (int ks = 1; ks <= 16; ks *= 2) {
        if (ktiles % ks || tiles * ks > 4096) continue;
        h->tune_role = PR_WGRAD; h->tune_layer = l; h->tune_cfg = c; h->tune_ks = ks;
        h->tune_wpart = tmp;
        float t = time_launches(h, [&] { launch_wgrad(h, l, GATE_ALWAYS, wt); }, reps);
        if (ks > 1) t += (float)((double)ks * ly.in * ly.out * 4.0 / 3.0e12 * 1e3);  // fold reads
        if (t < best) { best = t; bc = c; bks = ks; }
      }
    }
    ly.wg_cfg = bc;
    ly.wg_ks = bks;
    ly.wg_rt = ly.inp / gemm_cfg_tile_rows(bc);
    ly.wg_ct = ly.outp / gemm_cfg_tile_cols(bc);
  }
  h->tune_role = h->tune_layer = h->tune_cfg = -1;
  h->tune_wpart = nullptr;
  h->prof_mask = saved_mask;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  dfree(h, tmp);
  if (int rc = ensure_wpart(h)) return rc;
  h->tuned = true;
  h->tuned_data = true;
  if (h->fwd_graph) { (void)hipGraphExecDestroy(h->fwd_graph); h->fwd_graph = nullptr; }
  if (h->adj_graph) { (void)hipGraphExecDestroy(h->adj_graph); h->adj_graph = nullptr; }
  HIPCHK(h, hipGetLastError());
  return 0;
}

int maybe_autotune(node_handle* h) {
  if (!h->autotune || h->tuned || !h->have_params) return 0;
  return autotune_impl(h);
}

int ensure_adjoint(node_handle* h) {
  if (h->mu) return 0;
  const int64_t sv = (int64_t)h->Bp * h->D;
  const size_t act = (size_t)h->Bp * h->Dp * h->esz * h->planes;
  if (int rc = dalloc_t(h, &h->lam, sv * 4)) return rc;
  if (int rc = dalloc_t(h, &h->lam_new, sv * 4)) return rc;
  if (int rc = dalloc_t(h, &h->kl, 8 * sv * 4)) return rc;
  if (int rc = dalloc(h, &h->Lm, act)) return rc;
  if (int rc = dalloc(h, &h->LmT, act / h->Bp * h->ldT)) return rc;
  if (h->cnf) {
    const int L = h->L;
    const size_t unit = (size_t)h->Bp * h->esz * h->planes;
    if (int rc = dalloc(h, &h->E0, unit * h->layers[0].inp)) return rc;
    h->Tn.assign(L - 1, nullptr);
    h->Ab.assign(L, nullptr);
    h->src2.assign(L - 1, nullptr);
    h->E0s.assign(h->npr, nullptr);
    h->AbLs.assign(h->npr, nullptr);
    h->E0s[0] = h->E0;
    for (int k = 1; k < h->npr; ++k) {
      if (int rc = dalloc(h, &h->E0s[k], unit * h->layers[0].inp)) return rc;
      if (int rc = dalloc(h, &h->AbLs[k], unit * h->layers[L - 1].outp)) return rc;
    }
    for (int l = 0; l < L; ++l) {
      const int op = h->layers[l].outp;
      if (int rc = dalloc(h, &h->Ab[l], unit * op)) return rc;
      if (l == L - 1) { h->AbLs[0] = h->Ab[l]; break; }
      if (int rc = dalloc(h, &h->Tn[l], unit * op)) return rc;
      if (int rc = dalloc_t(h, &h->src2[l], (size_t)h->Bp * op * sizeof(float))) return rc;
    }
  }
  if (int rc = dalloc_t(h, &h->mu, 2 * h->P * 8)) return rc;
  if (int rc = dalloc_t(h, &h->accb, h->P * 4)) return rc;
  if (int rc = dalloc_t(h, &h->acce, h->P * 4)) return rc;
  if (int rc = dalloc_t(h, &h->gf0, h->P * 4)) return rc;
  if (int rc = dalloc_t(h, &h->gf1, h->P * 4)) return rc;
  for (int l = 0; l < h->L - 1; ++l) {
    Layer& ly = h->layers[l];
    if (int rc = dalloc_t(h, &ly.bpart, (size_t)(h->Bp / 32) * ly.outp * sizeof(float))) return rc;
  }
  if (int rc = ensure_wpart(h)) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

AdjBufs adj_bufs(node_handle* h) {
  AdjBufs a{};
  a.lam = h->lam; a.lam_new = h->lam_new; a.kl = h->kl; a.U = h->ucur;
  a.L = h->Lm; a.LT = h->LmT; a.X0 = h->X0; a.X0T = h->X0T;
  a.ldl = h->Dp; a.ldlt = h->ldT; a.ldx = h->Dp; a.ldxt = h->ldT;
  a.pl = a.px = (int64_t)h->Bp * h->Dp;
  a.plt = a.pxt = (int64_t)h->ldT * h->Dp;
  a.tape = h->tape; a.fwd_times = h->fwd_times; a.cb_times = h->save_times;
  a.dLdout = h->dLdout_dev;
  a.mu = h->mu; a.accb = h->accb; a.acce = h->acce; a.gf0 = h->gf0; a.gf1 = h->gf1;
  a.P = h->P;
  if (h->sense != NODE_SENSE_INTERPOLATING) {
    a.ubs = h->bs_u; a.ubs_new = h->bs_unew; a.ku = h->bs_ku; a.fsaved = h->out_dev;
    a.ckpt = h->sense == NODE_SENSE_BACKSOLVE;
  }
  return a;
}

void init_ctl(node_handle* h, Ctl& c, const node_solve_opts& o, double t0, double t1) {
  std::memset(&c, 0, sizeof(c));
  c.t = t0;
  c.tprev = t0;
  c.tdir = t1 > t0 ? 1.0 : -1.0;
  c.t_end = t1;
  const double span = std::fabs(t1 - t0);
  c.dtmax_abs = o.dtmax > 0 ? o.dtmax : span;
  auto spacing = [](double x) { x = std::fabs(x); return std::nextafter(x, INFINITY) - x; };
  c.dtmin = o.dtmin > 0 ? o.dtmin : std::max(spacing(t0), spacing(t1));
  c.abstol = o.abstol > 0 ? o.abstol : 1e-6;
  c.reltol = o.reltol > 0 ? o.reltol : 1e-3;
  c.adaptive = o.adaptive ? 1 : 0;
  c.dt_fixed = std::fabs(o.dt);
  c.has_dt_ = (!o.adaptive || o.dt != 0.0) ? 1 : 0;
  c.dt = c.has_dt_ ? c.tdir * std::fabs(o.dt) : 0.0;
  c.qold = 1e-4;
  c.q11 = 1.0;
  c.maxiters = o.maxiters > 0 ? o.maxiters : 100000;
  c.nf = 1;
  c.jump_idx = -1;
  (void)h;
}

// Step attempts to enqueue before the next poll: the span left over the
// current step size, at least the save (tstop) times still ahead, plus one for
// a rejection, in whole units of `unit` attempts (one captured graph), at most
// 16.  Attempts enqueued past the end of the solve return at once on the
// device but still cost their launches, so the chunk follows the estimate
// rather than growing blindly.
int64_t next_chunk(const Ctl& c, int64_t saves_left, int unit) {
  const double span = std::fabs(c.t_end - c.t), step = std::fabs(c.dt);
  const double est = step > 0.0 ? std::ceil(std::min(span / step, 64.0)) : 64.0;
  int64_t n = std::max<int64_t>((int64_t)est, saves_left) + 1;
  n = (n + unit - 1) / unit * unit;
  return std::min<int64_t>(std::max<int64_t>(n, unit), std::max(16, unit));
}

int poll_ctl(node_handle* h) {
  HIPCHK(h, hipMemcpyAsync(h->hctl, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int upload_save_times(node_handle* h, const std::vector<double>& ts) {
  if ((int64_t)ts.size() + 1 > h->save_cap) {
    dfree(h, h->save_times);
    h->save_cap = std::max<int64_t>(ts.size() + 1, 64);
    if (int rc = dalloc_t(h, &h->save_times, h->save_cap * sizeof(double))) return rc;
  }
  if (!ts.empty())
    HIPCHK(h, hipMemcpyAsync(h->save_times, ts.data(), ts.size() * sizeof(double),
                             hipMemcpyHostToDevice, h->stream));
  return 0;
}

int ensure_out(node_handle* h, int64_t n) {
  const int64_t need = std::max<int64_t>(n, 1) * h->B * h->D;
  if (need <= h->out_dev_cap) return 0;
  dfree(h, h->out_dev);
  h->out_dev = nullptr;
  if (int rc = dalloc_t(h, &h->out_dev, need * sizeof(float))) return rc;
  h->out_dev_cap = need;
  return 0;
}

// ---------------------------------------------------------------- forward
constexpr int64_t TAPE_RING = 4;  // ring slots of a forward that keeps checkpoints only

// The forward Tsit5 loop on the device from d_u0 ([B][D]) over o.t0 -> o.t1
// with o's tolerances and step options; the first n_save of the uploaded save
// times are saved into out_dev.  ring > 0: the tape entries form a ring of
// that many slots (no step record survives the solve).  The final control
// block is left in h->hctl.
int fwd_run(node_handle* h, const node_solve_opts& o, const float* d_u0, int64_t n_save,
            int64_t ring) {
  const StateGeom sg = geom(h);
  Ctl c;
  init_ctl(h, c, o, o.t0, o.t1);
  c.n_save = n_save;
  c.tape_ring = ring;
  // (a few entries to start; the loop grows the record as the steps need it)
  if (int rc = ensure_tape(h, ring > 0 ? ring : 4)) return rc;
  HIPCHK(h, hipMemcpyAsync(h->ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->fwd_times, &o.t0, sizeof(double), hipMemcpyHostToDevice, h->stream));
  double* p0 = h->part;
  double* px = h->part + SEC;
  const int nb = state_blocks(sg);
  const int64_t xpl = (int64_t)h->Bp * h->layers[0].inp;  // X0 plane stride
  launch_fwd_init(sg, d_u0, h->tape, h->X0, xpl, h->ctl, p0, nb, h->stream);
  rhs_eval(h, 0, SM_INIT_F0, GATE_ALWAYS, px);
  const int nx = rhs_last_tiles(h);
  const double nel = (double)h->B * h->D;
  if (!c.has_dt_) {
    launch_initdt_a(h->ctl, p0, nb, px, nx, nullptr, 0, nel, h->stream);
    launch_fwd_initdt_prep(sg, h->tape, h->X0, xpl, h->ctl, h->stream);
    rhs_eval(h, 0, SM_INIT_F1, GATE_ALWAYS, px);
  }
  launch_initdt_b(h->ctl, px, nx, nullptr, 0, nel, h->save_times, 0, h->fwd_times, h->stream);
  HIPCHK(h, hipGetLastError());

  auto attempt = [&]() {
    launch_fwd_step_begin(sg, h->tape, h->X0, xpl, h->save_times, h->out_dev, h->ctl, h->stream);
    for (int s = 1; s < 7; ++s) rhs_eval(h, s, SM_STEP, GATE_ACTIVE, px);
    launch_fwd_controller(h->ctl, px, nx, h->save_times, h->fwd_times, nel, h->stream);
  };
  const bool graphs = h->use_graphs && h->prof_mask == 0;
  const int ga = h->graph_attempts;
  int64_t chunk = graphs ? ga : 2, enqueued = 0;
  for (;;) {
    if (int rc = poll_ctl(h)) return rc;
    if (h->hctl->done) break;
    if (enqueued > c.maxiters + 64)
      return fail(h, NODE_ERR_STATE, "step controller did not terminate (internal error)");
    if (enqueued > 0)
      chunk = next_chunk(*h->hctl, std::max<int64_t>(0, h->hctl->n_save - h->hctl->save_next),
                         graphs ? ga : 1);
    bool eager = !graphs;
    if (ring == 0 && h->tape_budget > 0) {
      // a tape budget: enqueue no more attempts than the budget has entries
      // for (one by one when fewer than a captured graph's worth are left)
      const int64_t room = h->tape_budget / (8 * (int64_t)h->Bp * h->D * (int64_t)sizeof(float)) -
                           (h->hctl->n + 2);
      if (room < chunk) {
        chunk = std::max<int64_t>(room, 1);
        eager = true;
      }
    }
    enqueued += chunk;
    if (ring == 0)
      if (int rc = ensure_tape(h, h->hctl->n + chunk + 2)) return rc;
    if (!eager) {
      const void* sig[3] = {h->tape, h->save_times, h->out_dev};
      if (h->fwd_graph && std::memcmp(sig, h->fwd_graph_sig, sizeof(sig)) != 0) {
        (void)hipGraphExecDestroy(h->fwd_graph);
        h->fwd_graph = nullptr;
      }
      if (!h->fwd_graph) {
        hipGraph_t gr = nullptr;
        HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
        for (int k = 0; k < ga; ++k) attempt();
        HIPCHK(h, hipStreamEndCapture(h->stream, &gr));
        HIPCHK(h, hipGraphInstantiate(&h->fwd_graph, gr, nullptr, nullptr, 0));
        (void)hipGraphDestroy(gr);
        std::memcpy(h->fwd_graph_sig, sig, sizeof(sig));
      }
      for (int64_t k = 0; k < chunk; k += ga) HIPCHK(h, hipGraphLaunch(h->fwd_graph, h->stream));
    } else {
      for (int64_t k = 0; k < chunk; ++k) attempt();
    }
    HIPCHK(h, hipGetLastError());
  }
  // pending saves of the final step
  launch_fwd_step_begin(sg, h->tape, h->X0, xpl, h->save_times, h->out_dev, h->ctl, h->stream);
  return poll_ctl(h);
}

int forward_impl(node_handle* h, const node_solve_opts* o, const float* d_u0,
                 const double* saveat, int64_t n_saveat, node_stats* st) {
  if (!h->have_params) return fail(h, NODE_ERR_STATE, "parameters not set");
  if (!o) return fail(h, NODE_ERR_INVALID, "opts is NULL");
  if (int rc = maybe_autotune(h)) return rc;
  if (!(o->t1 != o->t0) || !std::isfinite(o->t0) || !std::isfinite(o->t1))
    return fail(h, NODE_ERR_INVALID, "tspan must be finite with t1 != t0");
  if (!o->adaptive && !(o->dt != 0.0)) return fail(h, NODE_ERR_INVALID, "fixed-step solve needs dt");
  if (o->save_everystep && n_saveat > 0)
    return fail(h, NODE_ERR_UNSUPPORTED, "save_everystep together with saveat is not supported");
  const bool ring = h->ckpt && h->sense == NODE_SENSE_INTERPOLATING && !h->cnf;
  if (ring && o->save_everystep)
    return fail(h, NODE_ERR_UNSUPPORTED,
                "InterpolatingAdjoint(checkpointing = true) with save_everystep is not supported "
                "(its checkpoints would be every step)");
  if (ring && !(o->t1 > o->t0))
    return fail(h, NODE_ERR_UNSUPPORTED, "checkpointing needs t1 > t0");
  const double tdir = o->t1 > o->t0 ? 1.0 : -1.0;
  std::vector<double> ts;
  if (o->save_start) ts.push_back(o->t0);
  for (int64_t i = 0; i < n_saveat; ++i) {
    const double s = saveat[i];
    if (!(tdir * s > tdir * o->t0 && tdir * s < tdir * o->t1))
      return fail(h, NODE_ERR_INVALID, "saveat entries must lie strictly inside tspan");
    if (i > 0 && !(tdir * s > tdir * saveat[i - 1]))
      return fail(h, NODE_ERR_INVALID, "saveat must be strictly monotone in the direction of time");
    ts.push_back(s);
  }
  if (o->save_end && !o->save_everystep) ts.push_back(o->t1);
  h->everystep = o->save_everystep != 0;
  h->fwd_ring = false;
  if (int rc = upload_save_times(h, ts)) return rc;
  if (int rc = ensure_out(h, (int64_t)ts.size())) return rc;
  const int64_t sv = (int64_t)h->B * h->D;
  if (ring) {
    if (!h->ck_u0) {
      if (int rc = dalloc_t(h, &h->ck_u0, sv * sizeof(float))) return rc;
      if (int rc = dalloc_t(h, &h->ck_uT, sv * sizeof(float))) return rc;
      if (int rc = dalloc_t(h, &h->ctl_save, sizeof(Ctl))) return rc;
    }
    HIPCHK(h, hipMemcpyAsync(h->ck_u0, d_u0, sv * sizeof(float), hipMemcpyDeviceToDevice,
                             h->stream));
  }
  if (int rc = fwd_run(h, *o, d_u0, h->everystep ? 0 : (int64_t)ts.size(), ring ? TAPE_RING : 0))
    return rc;
  const Ctl& r = *h->hctl;
  h->n_fwd = r.n;
  if (ring) {  // u(t1): the uprev of the final entry of the ring
    const int64_t ent = 8 * (int64_t)h->Bp * h->D;
    HIPCHK(h, hipMemcpyAsync(h->ck_uT, h->tape + (r.n % TAPE_RING) * ent, sv * sizeof(float),
                             hipMemcpyDeviceToDevice, h->stream));
  }
  const StateGeom sg = geom(h);
  if (h->everystep) {
    // every accepted step: states are tape[j].uprev, j = 0..n
    std::vector<double> ft(r.n + 1);
    HIPCHK(h, hipMemcpyAsync(ft.data(), h->fwd_times, (r.n + 1) * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    ts.clear();
    int64_t j0 = o->save_start ? 0 : 1;
    for (int64_t j = j0; j <= r.n; ++j) ts.push_back(ft[j]);
    if (int rc = ensure_out(h, (int64_t)ts.size())) return rc;
    launch_gather(sg, h->tape + j0 * sg.ent, sg.ent, h->out_dev, (int64_t)ts.size(), h->stream);
    if (int rc = upload_save_times(h, ts)) return rc;
    h->nsaved = (int64_t)ts.size();
  } else {
    h->nsaved = r.save_next;
  }
  ts.resize(h->nsaved);
  h->h_save = ts;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipGetLastError());
  node_stats s{};
  s.nf = r.nf; s.naccept = r.naccept; s.nreject = r.nreject; s.nsaved = h->nsaved;
  s.retcode = r.retcode; s.t_final = r.t;
  h->fwd_stats = s;
  h->fwd_opts = *o;
  h->fwd_ring = ring;
  h->have_forward = true;
  if (st) *st = s;
  return 0;
}

// ---------------------------------------------------------------- checkpoints
// Device state of checkpoint k ([B][D]).
const float* ck_state(const node_handle* h, size_t k) {
  const int64_t src = h->ck_src[k];
  if (src == -1) return h->ck_u0;
  if (src == -2) return h->ck_uT;
  return h->out_dev + src * (int64_t)h->B * h->D;
}

// Checkpoints of the last forward: t0, the saved times strictly inside the
// span, t1 (SciMLSensitivity's checkpoints = the solution's saved times, the
// span's ends added so that every adjoint stage time lies in an interval).
void ck_plan(node_handle* h) {
  const node_solve_opts& o = h->fwd_opts;
  h->ck_t.assign(1, o.t0);
  h->ck_src.assign(1, -1);
  for (size_t i = 0; i < h->h_save.size(); ++i)
    if (h->h_save[i] > o.t0 && h->h_save[i] < o.t1) {
      h->ck_t.push_back(h->h_save[i]);
      h->ck_src.push_back((int64_t)i);
    }
  h->ck_t.push_back(o.t1);
  h->ck_src.push_back(-2);
}

// Re-solve the forward over checkpoint interval k (ck_t[k] -> ck_t[k+1]) from
// its checkpoint state into the tape: the forward's tolerances and step mode,
// dtmax / dtmin of the interval problem, initial dt dt_init (0: the
// heuristic).  Then the state of checkpoint k + 1 goes to tape entry n + 1,
// where the resumed adjoint's FSAL re-evaluation at that time reads it.
int ck_resolve(node_handle* h, size_t k, double dt_init) {
  node_solve_opts o = h->fwd_opts;
  o.t0 = h->ck_t[k];
  o.t1 = h->ck_t[k + 1];
  o.dtmax = 0.0;
  o.dtmin = 0.0;
  if (o.adaptive) o.dt = dt_init;
  o.save_start = o.save_end = o.save_everystep = 0;
  if (int rc = fwd_run(h, o, ck_state(h, k), 0, 0)) return rc;
  const Ctl r = *h->hctl;
  if (r.retcode != NODE_RET_SUCCESS)
    return fail(h, NODE_ERR_STATE, "checkpoint re-solve of the forward interval [" +
                                       std::to_string(o.t0) + ", " + std::to_string(o.t1) +
                                       "] failed (retcode " + std::to_string(r.retcode) + ")");
  h->ck_n = r.n;
  h->ck_last_dt = std::fabs(r.t - r.tprev);  // span of the last step
  h->ck_nf += r.nf;
  if (int rc = ensure_tape(h, r.n + 2)) return rc;
  const int64_t ent = 8 * (int64_t)h->Bp * h->D;
  HIPCHK(h, hipMemcpyAsync(h->tape + (r.n + 1) * ent, ck_state(h, k + 1),
                           (int64_t)h->B * h->D * sizeof(float), hipMemcpyDeviceToDevice,
                           h->stream));
  return 0;
}

// ---------------------------------------------------------------- CNF
// One right-hand side of the continuous normalizing flow: hidden layers
// (keeping act'), the last layer f (plain, FP32), the probe's backprop
// e^T df/dx through the data-gradient GEMMs, and the fused stage GEMM whose
// epilogue forms e^T (df/dz) e and the Tsit5 stage of [z; log p].
void cnf_eval(node_handle* h, int stage, int mode, int gate, double* part) {
  const int L = h->L;
  for (int l = 0; l < L - 1; ++l) launch_hidden(h, l, gate, true);
  {  // f = last layer (identity activation), FP32
    const int l = L - 1;
    const Layer& ly = h->layers[l];
    GemmArgs g{};
    g.A = h->H[l - 1];
    g.lda = ly.inp; g.B = ly.Wr; g.ldb = ly.inp; g.K = ly.inp;
    g.pa = (int64_t)h->Bp * ly.inp; g.pb = (int64_t)ly.outp * ly.inp; g.ksplit = 1;
    GemmPlan p = plan_role(h, PR_FWD_STAGE, l, h->Bp, ly.outp, 1);
    g.row_tiles = p.rt; g.col_tiles = p.ct;
    EpiFwdStage e{};
    e.bias = ly.bias; e.act = ly.act; e.D = ly.out; e.B = h->B; e.stage = 0; e.mode = SM_PLAIN;
    e.scratch = h->fbuf;
    ProfScope ps(h, PR_FWD_STAGE, l);
    launch_gemm_fwdstage(h->prec, p.cfg, g, e, h->ctl, gate, h->stream);
  }
  // probe backprop(s): the Hutchinson probe, or each unit probe of the exact
  // trace in turn (the last pass forms the stage)
  const int npass = h->cnf_exact ? h->cnf_dm : 1;
  void* lm = h->Lm;
  for (int k = 0; k < npass; ++k) {
    if (h->cnf_exact) h->Lm = h->LmS[k];
    for (int l = L - 1; l >= 1; --l) launch_bwd_data(h, l, gate, false);
    const Layer& l0 = h->layers[0];
    GemmArgs g{};
    g.A = h->Dl[0];
    g.lda = l0.outp; g.B = l0.Wc; g.ldb = l0.outp; g.K = l0.outp;
    g.pa = (int64_t)h->Bp * l0.outp; g.pb = (int64_t)l0.inp * l0.outp; g.ksplit = 1;
    GemmPlan p = plan_role(h, PR_BWD_LAST, 0, h->Bp, l0.inp, 2);
    g.row_tiles = p.rt; g.col_tiles = p.ct;
    const StateGeom sg = geom(h);
    EpiCnfStage e{};
    e.D = h->D; e.B = h->B; e.dm = sg.dm; e.tcat = sg.tcat; e.stage = stage; e.mode = mode;
    e.tape = h->tape; e.ent = sg.ent; e.sv = sg.sv;
    e.X0 = h->X0; e.ldx = l0.inp; e.px = (int64_t)h->Bp * l0.inp;
    e.f = h->fbuf; e.e = h->probe; e.part = part;
    e.reg = h->cnf_reg ? 1 : 0; e.pass = k; e.npass = npass; e.trc = h->trc; e.wacc = h->wacc;
    ProfScope ps(h, PR_BWD_LAST, 0);
    launch_gemm_cnfstage(h->prec, p.cfg, g, e, h->ctl, gate, h->stream);
  }
  h->Lm = lm;
}

int cnf_last_tiles(node_handle* h) {
  GemmPlan p = plan_role(h, PR_BWD_LAST, 0, h->Bp, h->layers[0].inp, 2);
  return p.rt * p.ct;
}

// ---------------------------------------------------------------- backward
int backward_impl(node_handle* h, const float* d_dLdout, node_stats* st) {
  if (!h->have_forward) return fail(h, NODE_ERR_STATE, "no forward solve to differentiate");
  if (h->cnf && !h->cnf_adj) return fail(h, NODE_ERR_UNSUPPORTED, "CNF handle: use node_cnf_backward");
  if (h->layers[h->L - 1].act != NODE_ACT_IDENTITY)
    return fail(h, NODE_ERR_UNSUPPORTED, "adjoint: the last Dense layer must be linear");
  if (h->fwd_stats.retcode != NODE_RET_SUCCESS)
    return fail(h, NODE_ERR_STATE, "forward solve did not succeed (retcode != Success)");
  if (!(h->fwd_opts.t1 > h->fwd_opts.t0))
    return fail(h, NODE_ERR_UNSUPPORTED, "adjoint needs a forward solve with t1 > t0");
  if (int rc = ensure_adjoint(h)) return rc;
  if (h->sense != NODE_SENSE_INTERPOLATING && !h->bs_u) {
    const int64_t sv = (int64_t)h->Bp * h->D;
    if (int rc = dalloc_t(h, &h->bs_u, sv * 4)) return rc;
    if (int rc = dalloc_t(h, &h->bs_unew, sv * 4)) return rc;
    if (int rc = dalloc_t(h, &h->bs_ku, 8 * sv * 4)) return rc;
    if (int rc = dalloc_t(h, &h->bs_fu, sv * 4)) return rc;
  }
  if (int rc = maybe_autotune(h)) return rc;
  const bool ck = h->fwd_ring;
  if (ck && h->sense != NODE_SENSE_INTERPOLATING)
    return fail(h, NODE_ERR_STATE, "the last forward kept checkpoints only "
                                   "(InterpolatingAdjoint(checkpointing = true)); solve it again "
                                   "for another sensealg");
  if (!ck && h->ckpt && h->sense == NODE_SENSE_INTERPOLATING && !h->cnf)
    return fail(h, NODE_ERR_STATE, "checkpointing was switched on after the forward solve");
  const node_solve_opts& o = h->fwd_opts;
  const StateGeom sg = geom(h);
  const int64_t ncb = (int64_t)h->h_save.size();
  size_t kcur = 0;  // checkpoint interval on the tape
  h->ck_nf = 0;
  if (ck) {  // the last interval first; the adjoint pauses at its lower end
    ck_plan(h);
    kcur = h->ck_t.size() - 2;
    if (int rc = ck_resolve(h, kcur, 0.0)) return rc;
  }
  Ctl c;
  init_ctl(h, c, o, o.t1, o.t0);
  c.n_fwd = ck ? h->ck_n : h->n_fwd;
  c.n_save = ncb;
  if (ck && kcur > 0) {
    c.has_pause = 1;
    c.pause_t = h->ck_t[kcur];
  }
  int64_t next = ncb - 1;
  if (next >= 0 && h->h_save[next] == o.t1) { c.jump_idx = (int)next; --next; }
  c.save_next = next;
  HIPCHK(h, hipMemsetAsync(h->mu, 0, 2 * h->P * sizeof(double), h->stream));
  HIPCHK(h, hipMemcpyAsync(h->ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, h->stream));
  AdjBufs ab = adj_bufs(h);
  ab.dLdout = d_dLdout;
  double* p0 = h->part;
  double* px = h->part + SEC;
  double* pm = h->part + 2 * SEC;
  const int nb = state_blocks(sg);
  const int nmb = mu_blocks(h->P);
  const int nx = vjp_last_tiles(h);
  // error norm over [lambda; mu], or [u; lambda; mu] when u is integrated too
  const double nel = (h->sense != NODE_SENSE_INTERPOLATING ? 2.0 : 1.0) * h->B * h->D + (double)h->P;

  launch_set_stage_interp(h->ctl, 0, o.t1, h->fwd_times, h->stream);
  launch_adj_init(sg, ab, h->ctl, p0, h->stream);
  vjp_eval(h, 0, SM_INIT_F0, GATE_ALWAYS, WG_STORE_CUR, 0, 0, 0, 0, px, nullptr, nullptr);
  launch_mu_norm(1, ab, h->ctl, pm, nmb, GATE_ALWAYS, h->stream);
  if (!c.has_dt_) {
    launch_initdt_a(h->ctl, p0, nb, px, nx, pm, nmb, nel, h->stream);
    launch_set_stage_interp(h->ctl, 7, 0.0, h->fwd_times, h->stream);
    launch_adj_initdt_prep(sg, ab, h->ctl, h->stream);
    vjp_eval(h, 0, SM_INIT_F1, GATE_ALWAYS, WG_STORE_PTR, 0, 0, 0, 0, px, nullptr, h->accb);
    launch_mu_norm(2, ab, h->ctl, pm, nmb, GATE_ALWAYS, h->stream);
  }
  launch_initdt_b(h->ctl, px, nx, pm, nmb, nel, h->save_times, 1, h->fwd_times, h->stream);
  HIPCHK(h, hipGetLastError());

  auto attempt = [&]() {
    launch_adj_step_begin(sg, ab, h->ctl, GATE_ALWAYS, h->stream);
    vjp_eval(h, 0, SM_STEP, GATE_FSAL, WG_STORE_CUR, 0, 0, 0, 0, px, nullptr, nullptr);
    for (int s = 1; s < 7; ++s) {
      int fl;
      float cb = (float)kB[s], ce = (float)kBT[s], cb1 = 0, ce1 = 0;
      if (s == 1) { fl = WG_INIT; cb1 = (float)kB[0]; ce1 = (float)kBT[0]; }
      else if (s < 6) fl = WG_B | WG_E;
      else fl = WG_E | WG_STORE_NEXT;
      vjp_eval(h, s, SM_STEP, GATE_ACTIVE, fl, cb, ce, cb1, ce1, px, nullptr, nullptr);
    }
    launch_mu_norm(0, ab, h->ctl, pm, nmb, GATE_ACTIVE, h->stream);
    launch_adj_controller(h->ctl, px, nx, pm, nmb, nel, h->save_times, h->fwd_times, h->stream);
  };
  const bool graphs = h->use_graphs && h->prof_mask == 0;
  const int ga = h->graph_attempts;
  int64_t chunk = graphs ? ga : 2, enqueued = 0;
  bool resumed = false;
  for (;;) {
    if (int rc = poll_ctl(h)) return rc;
    if (h->hctl->done) {
      if (!(ck && h->hctl->paused)) break;
      // checkpoint reached: re-solve the forward interval below it, resume
      HIPCHK(h, hipMemcpyAsync(h->ctl_save, h->ctl, sizeof(Ctl), hipMemcpyDeviceToDevice,
                               h->stream));
      --kcur;
      if (int rc = ck_resolve(h, kcur, h->ck_last_dt)) return rc;
      // a re-solve may have grown (moved) the tape: the attempts' buffer
      // table follows it (the captured graph is re-captured below)
      ab = adj_bufs(h);
      ab.dLdout = d_dLdout;
      launch_adj_resume(h->ctl, h->ctl_save, h->ck_n, kcur > 0 ? h->ck_t[kcur] : 0.0,
                        kcur > 0 ? 1 : 0, h->fwd_times, h->stream);
      HIPCHK(h, hipGetLastError());
      resumed = true;
      continue;
    }
    if (enqueued > c.maxiters + 64)
      return fail(h, NODE_ERR_STATE, "adjoint step controller did not terminate (internal error)");
    if (enqueued > 0 && !resumed)
      chunk = next_chunk(*h->hctl, h->hctl->save_next + 1, graphs ? ga : 1);
    else if (resumed)
      chunk = graphs ? ga : 2;
    resumed = false;
    enqueued += chunk;
    if (graphs) {
      // the tape may have moved (a checkpoint re-solve grew it)
      const void* asig[4] = {h->tape, h->save_times, d_dLdout, h->out_dev};
      if (h->adj_graph && std::memcmp(asig, h->adj_graph_sig, sizeof(asig)) != 0) {
        (void)hipGraphExecDestroy(h->adj_graph);
        h->adj_graph = nullptr;
      }
      if (!h->adj_graph) {
        hipGraph_t gr = nullptr;
        HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
        for (int k = 0; k < ga; ++k) attempt();
        HIPCHK(h, hipStreamEndCapture(h->stream, &gr));
        HIPCHK(h, hipGraphInstantiate(&h->adj_graph, gr, nullptr, nullptr, 0));
        (void)hipGraphDestroy(gr);
        std::memcpy(h->adj_graph_sig, asig, sizeof(asig));
      }
      for (int64_t k = 0; k < chunk; k += ga) HIPCHK(h, hipGraphLaunch(h->adj_graph, h->stream));
    } else {
      for (int64_t k = 0; k < chunk; ++k) attempt();
    }
    HIPCHK(h, hipGetLastError());
  }
  launch_adj_step_begin(sg, ab, h->ctl, GATE_ALWAYS, h->stream);
  if (int rc = poll_ctl(h)) return rc;
  HIPCHK(h, hipGetLastError());
  const Ctl& r = *h->hctl;
  if (st) {
    node_stats s{};
    s.nf = r.nf; s.naccept = r.naccept; s.nreject = r.nreject; s.nsaved = 0;
    s.retcode = r.retcode; s.t_final = r.t;
    *st = s;
  }
  return 0;
}

// ---------------------------------------------------------------- ragged batches
// LDS layout of the per-trajectory kernels: per layer W rows [16*to][16*ti + 4]
// then the bias [16*to]
int ragged_model(node_handle* h, RaggedModel& m, std::string& why) {
  const node_desc& d = h->desc;
  m = RaggedModel{};
  if (d.n_layers > RG_MAXL) { why = "per-trajectory solves support at most 4 Dense layers"; return -1; }
  if (d.widths[0] > RG_MAXD) { why = "per-trajectory solves support state dimension D <= 32"; return -1; }
  m.L = d.n_layers;
  m.map = d.input_map;
  int off = 0;
  for (int l = 0; l < m.L; ++l) {
    const int in = d.widths[l], out = d.widths[l + 1];
    if (l < m.L - 1 && out > 16 * RG_HT) { why = "per-trajectory solves support hidden widths <= 128"; return -1; }
    m.in[l] = in;
    m.out[l] = out;
    m.act[l] = d.acts[l];
    m.ti[l] = (in + 15) / 16;
    m.to[l] = (out + 15) / 16;
    m.pitch[l] = 16 * m.ti[l] + 4;
    m.lw[l] = off;
    off += 16 * m.to[l] * m.pitch[l];
    m.lb[l] = off;
    off += 16 * m.to[l];
    m.w0[l] = h->layers[l].w0;
    m.b0[l] = h->layers[l].b0;
  }
  // 128 zero floats after the last layer: a sweep reads whole 8-tile rows
  // (k-tiles past a layer's own are read but not used), which may run up to
  // 128 floats past the last layer's weights
  off += 128;
  m.lds_floats = off;
  if ((size_t)off * sizeof(float) > 160 * 1024) { why = "model weights exceed the 160 KB LDS of a workgroup"; return -1; }
  int so = 0;
  for (int l = 0; l < m.L; ++l) {
    m.sh[l] = so;
    so += 16 * m.ti[l];
    m.sd[l] = so;
    so += 16 * m.to[l];
  }
  m.sv = so;
  m.P = h->P;
  m.p = h->pflat;
  // rows of rw tiles: the longest tile dimension of any layer, rounded up to
  // a power of two (small models stream no 8-tile padding)
  int tmax = 1;
  for (int l = 0; l < m.L; ++l) tmax = std::max(tmax, std::max(m.ti[l], m.to[l]));
  m.rw = tmax <= 1 ? 1 : tmax <= 2 ? 2 : tmax <= 4 ? 4 : RG_HT;
  int64_t no = 0;
  for (int l = 0; l < m.L; ++l) {
    m.rowout[l] = m.to[l] >= m.ti[l];
    m.nrow[l] = m.rowout[l] ? m.ti[l] : m.to[l];
    m.nw0[l] = no;
    no += (int64_t)m.nrow[l] * m.rw * 256;
  }
  m.ntile = (int)(no / 256);
  for (int l = 0; l < m.L; ++l) {
    m.nb0[l] = no;
    no += m.out[l];
  }
  m.Pn = (no + 63) / 64 * 64;
  return 0;
}

template <typename P>
int grow(node_handle* h, P** p, int64_t need, int64_t& cap_elems, int64_t new_cap) {
  if (*p && need <= cap_elems) return 0;
  dfree(h, *p);
  *p = nullptr;
  cap_elems = new_cap;
  return dalloc_t(h, p, (size_t)new_cap * sizeof(P));
}

int ragged_grid(node_handle* h, int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(h->n_cu, (n + 63) / 64));
}

int ragged_forward_impl(node_handle* h, const node_solve_opts* o, int64_t n, const double* t1,
                        const float* u0, const int64_t* soff, const double* st, float* out,
                        int32_t* ret, int32_t* tstats, node_ragged_stats* stats) {
  if (!o || !t1 || !u0 || !soff || n < 1) return fail(h, NODE_ERR_INVALID, "null argument or n < 1");
  if (!h->have_params) return fail(h, NODE_ERR_STATE, "parameters not set");
  RaggedModel m;
  std::string why;
  if (ragged_model(h, m, why)) return fail(h, NODE_ERR_UNSUPPORTED, why);
  if (soff[0] != 0) return fail(h, NODE_ERR_INVALID, "save_offsets[0] must be 0");
  if (!o->adaptive && !(o->dt > 0)) return fail(h, NODE_ERR_INVALID, "fixed-step solve needs dt > 0");
  for (int64_t i = 0; i < n; ++i) {
    if (!(t1[i] > o->t0)) return fail(h, NODE_ERR_INVALID, "every trajectory needs t1 > t0");
    if (soff[i + 1] < soff[i]) return fail(h, NODE_ERR_INVALID, "save_offsets must be non-decreasing");
    for (int64_t k = soff[i]; k < soff[i + 1]; ++k) {
      if (!(st[k] >= o->t0 && st[k] <= t1[i])) return fail(h, NODE_ERR_INVALID, "save time outside [t0, t1]");
      if (k > soff[i] && !(st[k] > st[k - 1])) return fail(h, NODE_ERR_INVALID, "save times must increase");
    }
  }
  const int64_t ns = soff[n];
  if (ns > 0 && (!st || !out)) return fail(h, NODE_ERR_INVALID, "save times / output missing");
  const int D = h->D, Dp = (D + 15) / 16 * 16;
  auto& r = h->rg;
  r.have = false;
  int64_t c1 = r.n_cap, c2 = r.n_cap, c3 = r.n_cap, c4 = r.n_cap, c5 = r.n_cap, c6 = r.n_cap;
  if (!r.t1 || n > r.n_cap) {
    const int64_t cap = std::max<int64_t>(n, 64);
    if (grow(h, &r.t1, n, c1, cap) || grow(h, &r.u0, n * D, c2, cap * D) ||
        grow(h, &r.soff, n + 1, c3, cap + 1) || grow(h, &r.ret, n, c4, cap) ||
        grow(h, &r.tstats, 4 * n, c5, 4 * cap) || grow(h, &r.last, n, c6, cap))
      return NODE_ERR_HIP;
    r.n_cap = cap;
  }
  if (!r.st || ns > r.ns_cap) {
    const int64_t cap = std::max<int64_t>(ns, 64);
    int64_t a1 = r.ns_cap, a2 = r.ns_cap;
    if (grow(h, &r.st, ns, a1, cap) || grow(h, &r.out, ns * D, a2, cap * D)) return NODE_ERR_HIP;
    r.ns_cap = cap;
  }
  if (!r.ctr) {
    if (dalloc_t(h, &r.ctr, 4 * sizeof(int)) || dalloc_t(h, &r.stats, 4 * sizeof(unsigned long long)))
      return NODE_ERR_HIP;
  }
  if (h->n_cu == 0) {
    int dev = 0;
    HIPCHK(h, hipGetDevice(&dev));
    HIPCHK(h, hipDeviceGetAttribute(&h->n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  }
  HIPCHK(h, hipMemcpyAsync(r.t1, t1, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(r.u0, u0, n * D * sizeof(float), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(r.soff, soff, (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, h->stream));
  if (ns > 0)
    HIPCHK(h, hipMemcpyAsync(r.st, st, ns * sizeof(double), hipMemcpyHostToDevice, h->stream));
  if (r.rec_cap < n * 16) {
    const int64_t cap = n * 64;
    dfree(h, r.rec_t); dfree(h, r.rec_prev); dfree(h, r.rec_data);
    r.rec_t = nullptr; r.rec_prev = nullptr; r.rec_data = nullptr;
    if (dalloc_t(h, &r.rec_t, cap * sizeof(double)) || dalloc_t(h, &r.rec_prev, cap * sizeof(int)) ||
        dalloc_t(h, &r.rec_data, (size_t)cap * 8 * Dp * sizeof(float)))
      return NODE_ERR_HIP;
    r.rec_cap = cap;
  }
  RaggedFwdArgs a{};
  a.m = m;
  a.D = D;
  a.Dp = Dp;
  a.n = n;
  a.t1 = r.t1; a.u0 = r.u0; a.save_off = r.soff; a.save_t = r.st; a.out = r.out;
  a.retcode = r.ret; a.traj_stats = r.tstats; a.last_rec = r.last;
  a.queue = r.ctr; a.rec_count = r.ctr + 1;
  a.t0 = o->t0;
  a.abstol = o->abstol > 0 ? o->abstol : 1e-6;
  a.reltol = o->reltol > 0 ? o->reltol : 1e-3;
  a.dtmax = o->dtmax;
  a.dtmin = o->dtmin;
  a.dt_init = std::fabs(o->dt);
  a.maxiters = o->maxiters > 0 ? o->maxiters : 100000;
  a.adaptive = o->adaptive ? 1 : 0;
  a.stats = r.stats;
  const int grid = ragged_grid(h, n);
  int ctr[4];
  for (int attempt = 0; attempt < 3; ++attempt) {
    a.rec_t = r.rec_t; a.rec_prev = r.rec_prev; a.rec_data = r.rec_data; a.rec_cap = r.rec_cap;
    HIPCHK(h, hipMemsetAsync(r.ctr, 0, 4 * sizeof(int), h->stream));
    HIPCHK(h, hipMemsetAsync(r.stats, 0, 4 * sizeof(unsigned long long), h->stream));
    if (ns > 0) HIPCHK(h, hipMemsetAsync(r.out, 0, ns * D * sizeof(float), h->stream));
    launch_ragged_fwd(a, grid, h->stream);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(ctr, r.ctr, sizeof(ctr), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (ctr[1] <= r.rec_cap) break;
    // tape pool too small: grow to what the solve needed and run again
    const int64_t cap = (int64_t)ctr[1] + n;
    dfree(h, r.rec_t); dfree(h, r.rec_prev); dfree(h, r.rec_data);
    r.rec_t = nullptr; r.rec_prev = nullptr; r.rec_data = nullptr;
    if (dalloc_t(h, &r.rec_t, cap * sizeof(double)) || dalloc_t(h, &r.rec_prev, cap * sizeof(int)) ||
        dalloc_t(h, &r.rec_data, (size_t)cap * 8 * Dp * sizeof(float)))
      return NODE_ERR_HIP;
    r.rec_cap = cap;
  }
  if (ctr[1] > r.rec_cap) return fail(h, NODE_ERR_CAPACITY, "tape pool kept overflowing");
  r.h_ret.resize(n);
  HIPCHK(h, hipMemcpyAsync(r.h_ret.data(), r.ret, n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  if (ns > 0) HIPCHK(h, hipMemcpyAsync(out, r.out, ns * D * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (tstats) HIPCHK(h, hipMemcpyAsync(tstats, r.tstats, 4 * n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  unsigned long long hs[4];
  HIPCHK(h, hipMemcpyAsync(hs, r.stats, sizeof(hs), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (ret) std::memcpy(ret, r.h_ret.data(), n * sizeof(int));
  if (stats) {
    node_ragged_stats s{};
    s.nf = (int64_t)hs[0]; s.naccept = (int64_t)hs[1]; s.nreject = (int64_t)hs[2];
    s.records = ctr[1];
    for (int v : r.h_ret) s.n_success += v == NODE_RET_SUCCESS;
    *stats = s;
  }
  r.n = n;
  r.ns = ns;
  r.opts = *o;
  r.have = true;
  return 0;
}

int ragged_backward_impl(node_handle* h, const float* dLdout, float* dLdu0, double* dLdp,
                         float* dLdp_traj, int32_t* ret, int32_t* tstats,
                         node_ragged_stats* stats) {
  auto& r = h->rg;
  auto& b = h->ra;
  if (!r.have) return fail(h, NODE_ERR_STATE, "no per-trajectory forward solve to differentiate");
  for (int v : r.h_ret)
    if (v != NODE_RET_SUCCESS)
      return fail(h, NODE_ERR_STATE, "a trajectory's forward solve did not succeed (retcode != Success)");
  if (r.ns > 0 && !dLdout) return fail(h, NODE_ERR_INVALID, "dLdout is NULL");
  RaggedModel m;
  std::string why;
  if (ragged_model(h, m, why)) return fail(h, NODE_ERR_UNSUPPORTED, why);
  const int64_t n = r.n, ns = r.ns, P = h->P, Pn = m.Pn;
  const int D = h->D, Dp = (D + 15) / 16 * 16;
  // A workgroup per CU (the weights fill its LDS) of up to 4 waves, one per
  // SIMD.  A wave's time is its VJP rounds (as many as its longest slot
  // needs) plus one parameter pass per step of every slot, each pass bound
  // by the wave's own memory latency: so the trajectories are spread over as
  // many waves as the SIMDs hold (spw slots of 16 active per wave), and all
  // of them start at once when they fit.
  const int64_t max_waves = 4 * (int64_t)h->n_cu;
  const int spw = (int)std::min<int64_t>(RG_SLOTS, std::max<int64_t>(1, (n + max_waves - 1) / max_waves));
  const int64_t waves_all = (n + spw - 1) / spw;
  const int waves = (int)std::min<int64_t>(4, std::max<int64_t>(1, (waves_all + h->n_cu - 1) / h->n_cu));
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(h->n_cu, (waves_all + waves - 1) / waves));
  const int64_t slots = (int64_t)grid * waves * RG_SLOTS;      // stage blocks, row buffers
  const int64_t nslots = (int64_t)grid * waves * spw;          // active slots
  // act' of the VJPs in LDS beside the weights when every wave's area fits
  int act_tiles = 0;
  for (int l = 0; l < m.L; ++l) {
    m.act_off[l] = act_tiles;
    act_tiles += m.to[l];
  }
  m.act_stride = act_tiles * 256;
  if (((int64_t)m.lds_floats + (int64_t)waves * m.act_stride) * (int64_t)sizeof(float) > 160 * 1024)
    m.act_stride = 0;
  // start order (a schedule only: each trajectory's arithmetic is the same
  // in any slot).  A wave runs until its last slot's trajectory is done and
  // does one parameter pass per step of every slot, so the waves' sums of
  // steps should match.  A trajectory's adjoint steps are estimated as the
  // steps of the solve before it (the forward) plus its save times (the
  // adjoint stops at each of them); the trajectories, costliest first, are
  // dealt over the waves back and forth (wave w: ranks w, 2W-1-w, 2W+w, ...);
  // any beyond the launch's slots follow from the queue in the same order.
  const int64_t W = (int64_t)grid * waves, norder = std::max<int64_t>(n, nslots);
  std::vector<int> hts(4 * n), ord(norder, -1);
  std::vector<int64_t> hso(n + 1), rank(n), cost(n);
  HIPCHK(h, hipMemcpyAsync(hts.data(), r.tstats, 4 * n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(hso.data(), r.soff, (n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int64_t i = 0; i < n; ++i) {
    rank[i] = i;
    cost[i] = (int64_t)hts[4 * i + 1] + hts[4 * i + 2] + (hso[i + 1] - hso[i]);
  }
  std::stable_sort(rank.begin(), rank.end(), [&](int64_t x, int64_t y) { return cost[x] > cost[y]; });
  for (int64_t k = 0; k < n; ++k) {
    int64_t pos = k;
    if (k < nslots) {
      const int64_t lap = k / W, w = (lap & 1) ? W - 1 - k % W : k % W;
      pos = w * spw + lap;
    }
    ord[pos] = (int)rank[k];
  }
  if (grow(h, &b.order, norder, b.order_cap, norder) ||
      grow(h, &b.dLdout, std::max<int64_t>(ns * D, 1), b.dLdout_cap, std::max<int64_t>(ns * D, 64)) ||
      grow(h, &b.dLdu0, n * D, b.dLdu0_cap, n * D) ||
      grow(h, &b.rows, n * Pn, b.rows_cap, n * Pn) ||
      grow(h, &b.tmp, slots * Pn, b.tmp_cap, slots * Pn) ||
      grow(h, &b.sbuf, (slots + 1) * 8 * m.sv, b.sbuf_cap, (slots + 1) * 8 * m.sv) ||
      grow(h, &b.dp, P, b.dp_cap, P) ||
      grow(h, &b.part, RG_SUM_CHUNKS * P, b.part_cap, RG_SUM_CHUNKS * P))
    return NODE_ERR_HIP;
  if (ns > 0)
    HIPCHK(h, hipMemcpyAsync(b.dLdout, dLdout, ns * D * sizeof(float), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(b.order, ord.data(), norder * sizeof(int), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemsetAsync(b.rows, 0, (size_t)n * Pn * sizeof(float), h->stream));
  HIPCHK(h, hipMemsetAsync(r.ctr, 0, 4 * sizeof(int), h->stream));
  HIPCHK(h, hipMemsetAsync(r.stats, 0, 4 * sizeof(unsigned long long), h->stream));
  RaggedAdjArgs a{};
  a.m = m;
  a.D = D;
  a.Dp = Dp;
  a.n = n;
  a.t1 = r.t1; a.save_off = r.soff; a.save_t = r.st; a.dLdout = b.dLdout;
  a.last_rec = r.last; a.rec_t = r.rec_t; a.rec_prev = r.rec_prev; a.rec_data = r.rec_data;
  a.dLdu0 = b.dLdu0; a.dLdp_traj = b.rows; a.mu_tmp = b.tmp; a.sbuf = b.sbuf;
  a.queue = r.ctr; a.retcode = r.ret; a.traj_stats = r.tstats;
  a.order = b.order; a.norder = norder; a.nslots = nslots; a.spw = spw; a.nslots_all = slots;
  a.vjp_bf16 = h->rg_vjp_prec == NODE_PREC_BF16;
  const node_solve_opts* o = &r.opts;
  a.t0 = o->t0;
  a.abstol = o->abstol > 0 ? o->abstol : 1e-6;
  a.reltol = o->reltol > 0 ? o->reltol : 1e-3;
  a.dtmax = o->dtmax;
  a.dtmin = o->dtmin;
  a.dt_init = std::fabs(o->dt);
  a.maxiters = o->maxiters > 0 ? o->maxiters : 100000;
  a.adaptive = o->adaptive ? 1 : 0;
  a.stats = r.stats;
  launch_ragged_adj(a, grid, waves, h->stream);
  HIPCHK(h, hipGetLastError());
  launch_ragged_sum(b.rows, n, P, Pn, b.part, b.dp, h->stream);
  HIPCHK(h, hipGetLastError());
  std::vector<int> hret(n);
  unsigned long long hs[4];
  HIPCHK(h, hipMemcpyAsync(hret.data(), r.ret, n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(hs, r.stats, sizeof(hs), hipMemcpyDeviceToHost, h->stream));
  if (dLdu0) HIPCHK(h, hipMemcpyAsync(dLdu0, b.dLdu0, n * D * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (dLdp) HIPCHK(h, hipMemcpyAsync(dLdp, b.dp, P * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  if (dLdp_traj)
    HIPCHK(h, hipMemcpy2DAsync(dLdp_traj, P * sizeof(float), b.rows, Pn * sizeof(float),
                               P * sizeof(float), n, hipMemcpyDeviceToHost, h->stream));
  if (tstats) HIPCHK(h, hipMemcpyAsync(tstats, r.tstats, 4 * n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (ret) std::memcpy(ret, hret.data(), n * sizeof(int));
  if (stats) {
    node_ragged_stats s{};
    s.nf = (int64_t)hs[0]; s.naccept = (int64_t)hs[1]; s.nreject = (int64_t)hs[2];
    for (int v : hret) s.n_success += v == NODE_RET_SUCCESS;
    *stats = s;
  }
  return 0;
}

int copy_grads(node_handle* h, float* dLdu0, double* dLdp, hipMemcpyKind kind) {
  // lambda real rows are the first B*D floats (row stride D)
  if (dLdu0)
    HIPCHK(h, hipMemcpyAsync(dLdu0, h->lam, (size_t)h->B * h->D * sizeof(float), kind, h->stream));
  if (dLdp) {
    const double* mu = h->hctl->gsel_mu_ ? h->mu + h->P : h->mu;
    HIPCHK(h, hipMemcpyAsync(dLdp, mu, (size_t)h->P * sizeof(double), kind, h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

// ---------------------------------------------------------------- no-alloc RHS
// The fixed device array of node_rhs_device_noalloc: the model input, two
// ping-pong hidden-activation buffers and the output of one right-hand side
// live here instead of in memory allocated per handle or per call.  Each GPU
// holds its own copy until the process ends; host code uses the per-device
// address from hipGetSymbolAddress (the array's name is a host-side stand-in).
constexpr size_t NOALLOC_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEV = 64;
std::mutex g_scratch_mtx;
void* g_scratch_ptr[NOALLOC_MAX_DEV] = {};

// Device address of the current device's array, resolved once per device
// (node_create: the library's set-up, with the handle's device current).
hipError_t scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= NOALLOC_MAX_DEV) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_scratch_mtx);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace

// ================================================================ C ABI
extern "C" {

int32_t node_abi_version(void) { return NODE_HIP_ABI_VERSION; }

int64_t node_num_params(const node_desc* d) {
  std::string m;
  if (validate_desc(d, m)) return -1;
  int64_t p = 0;
  for (int l = 0; l < d->n_layers; ++l)
    p += (int64_t)d->widths[l] * d->widths[l + 1] + d->widths[l + 1];
  return p;
}

const char* node_last_error(const node_handle_t* h) { return h ? h->err.c_str() : "null handle"; }

int32_t node_create(const node_desc* desc, node_handle_t** out) {
  if (!out) return NODE_ERR_INVALID;
  *out = nullptr;
  std::string msg;
  if (int rc = validate_desc(desc, msg)) { std::fprintf(stderr, "node_create: %s\n", msg.c_str()); return rc; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    std::fprintf(stderr, "node_create: no HIP device\n");
    return NODE_ERR_HIP;
  }
  node_handle* h = new node_handle();
  for (auto& row : h->tcfg)
    for (int& c : row) c = -1;
  for (auto& row : h->tks)
    for (int& c : row) c = 1;
  h->desc = *desc;
  h->cnf_reg = (desc->input_map & NODE_MAP_CNF_REGULARIZE) != 0;
  h->cnf_exact = (desc->input_map & NODE_MAP_CNF_EXACT_TRACE) != 0;
  h->desc.input_map &= ~(NODE_MAP_CNF_REGULARIZE | NODE_MAP_CNF_EXACT_TRACE);
  h->L = desc->n_layers;
  h->cnf = h->desc.input_map == NODE_MAP_CNF || h->desc.input_map == NODE_MAP_CNF_TIME;
  h->cnf_dm = desc->widths[desc->n_layers];
  // ODE state width: [z; log p] (+ r1, r2 with regularize) for a CNF
  h->D = h->cnf ? h->cnf_dm + (h->cnf_reg ? 3 : 1) : desc->widths[0];
  h->npr = h->cnf && h->cnf_exact ? h->cnf_dm : 1;
  h->Dp = (int)rup(h->D, KPAD);
  h->B = (int)desc->batch;
  h->Bp = (int)rup(h->B, BPAD);
  // transposed operands: primal rows, then one slice per tangent chain (CNF)
  h->ldT = h->cnf ? (1 + h->npr) * h->Bp : h->Bp;
  h->prec = desc->precision;
  h->wg_prec = h->prec;
  h->esz = h->prec == NODE_PREC_FP32 ? 4 : 2;
  h->planes = h->prec == NODE_PREC_BF16X3 ? 2 : 1;
  h->P = node_num_params(desc);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return NODE_ERR_HIP;
  }
  if (upload_tableau(h->stream) != hipSuccess) {
    (void)hipStreamDestroy(h->stream);
    delete h;
    return NODE_ERR_HIP;
  }
  if (hipError_t e = scratch_address(&h->na_scratch); e != hipSuccess) {
    std::fprintf(stderr, "node_create: hipGetSymbolAddress(scratch): %s\n", hipGetErrorString(e));
    (void)hipStreamDestroy(h->stream);
    delete h;
    return NODE_ERR_HIP;
  }
  int64_t off = 0;
  int rc = 0;
  for (int l = 0; l < h->L && !rc; ++l) {
    Layer ly;
    ly.in = desc->widths[l];
    ly.out = desc->widths[l + 1];
    ly.inp = (int)rup(ly.in, KPAD);
    ly.outp = (int)rup(ly.out, KPAD);
    ly.act = desc->acts[l];
    ly.w0 = off;
    ly.b0 = off + (int64_t)ly.in * ly.out;
    off = ly.b0 + ly.out;
    rc = dalloc(h, &ly.Wr, (size_t)ly.outp * ly.inp * h->esz * h->planes);
    if (!rc) rc = dalloc(h, &ly.Wc, (size_t)ly.outp * ly.inp * h->esz * h->planes);
    {
      // weight-gradient plan: narrow layers get split-K so that the grid fills the chip
      GemmPlan p = plan(ly.inp, ly.outp);
      ly.wg_cfg = p.cfg; ly.wg_rt = p.rt; ly.wg_ct = p.ct; ly.wg_ks = 1;
      const int ktiles = (int)((int64_t)h->Bp * h->esz / 128);
      while (p.rt * p.ct * ly.wg_ks < 256 && ly.wg_ks < 16 && ktiles % (2 * ly.wg_ks) == 0)
        ly.wg_ks *= 2;
      ly.wg_ks0 = ly.wg_ks;
    }
    if (!rc) rc = dalloc_t(h, &ly.bias, (size_t)ly.outp * sizeof(float));
    h->layers.push_back(ly);
  }
  const size_t x0b = (size_t)h->Bp * h->Dp * h->esz * h->planes;
  if (!rc) rc = dalloc(h, &h->X0, x0b);
  if (!rc) rc = dalloc(h, &h->X0T, x0b / h->Bp * h->ldT);
  for (int l = 0; l < h->L - 1 && !rc; ++l) {
    const size_t hb = (size_t)h->Bp * h->layers[l].outp * h->esz * h->planes;
    const size_t hbt = hb / h->Bp * h->ldT;
    void *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr, *e = nullptr;
    rc = dalloc(h, &a, hb);
    if (!rc) rc = dalloc(h, &b, hbt);
    if (!rc) rc = dalloc(h, &c, hb);
    if (!rc) rc = dalloc(h, &d, hb);
    if (!rc) rc = dalloc(h, &e, hbt);
    h->H.push_back(a); h->HT.push_back(b); h->S.push_back(c); h->Dl.push_back(d); h->DlT.push_back(e);
  }
  if (!rc) rc = dalloc_t(h, &h->ctl, sizeof(Ctl));
  if (!rc) rc = dalloc_t(h, &h->part, 4 * SEC * sizeof(double));
  if (!rc) rc = dalloc_t(h, &h->scratch, (size_t)h->Bp * h->D * sizeof(float));
  if (!rc) rc = dalloc_t(h, &h->ucur, (size_t)h->Bp * h->D * sizeof(float));
  // split-K fixup workspace: up to 8 splits of a Bp x Dp (or last-layer) tile grid
  if (!rc) rc = dalloc_t(h, &h->fix_ws, (size_t)8 * h->Bp * std::max(h->Dp, h->layers[h->L - 1].outp) * sizeof(float));
  if (!rc) rc = dalloc_t(h, &h->fix_cnt, (size_t)(h->Bp / 32) * (std::max(h->Dp, h->layers[h->L - 1].outp) / 64) * sizeof(int));
  if (!rc && hipHostMalloc((void**)&h->hctl, sizeof(Ctl), 0) != hipSuccess) rc = NODE_ERR_HIP;
  if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = NODE_ERR_HIP;
  if (rc) {
    std::fprintf(stderr, "node_create: %s\n", h->err.c_str());
    node_destroy(h);
    return rc;
  }
  *out = h;
  return 0;
}

int32_t node_destroy(node_handle_t* h) {
  if (!h) return 0;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->fwd_graph) (void)hipGraphExecDestroy(h->fwd_graph);
  if (h->adj_graph) (void)hipGraphExecDestroy(h->adj_graph);
  for (void* p : h->allocs) (void)hipFree(p);
  for (auto& r : h->prof) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
  if (h->hctl) (void)hipHostFree(h->hctl);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

static int set_params_host(node_handle* h, const float* p, int64_t n) {
  if (!p) return fail(h, NODE_ERR_INVALID, "params is NULL");
  if (n != h->P) return fail(h, NODE_ERR_INVALID, "parameter count mismatch: expected " + std::to_string(h->P));
  for (const Layer& ly : h->layers) {
    const size_t cnt = (size_t)ly.outp * ly.inp;
    std::vector<float> wr(cnt, 0.f), wc(cnt, 0.f), b(ly.outp, 0.f);
    for (int i = 0; i < ly.in; ++i)
      for (int o = 0; o < ly.out; ++o) {
        const float w = p[ly.w0 + (int64_t)i * ly.out + o];  // Julia W[o, i]
        wr[(size_t)o * ly.inp + i] = w;
        wc[(size_t)i * ly.outp + o] = w;
      }
    for (int o = 0; o < ly.out; ++o) b[o] = p[ly.b0 + o];
    if (h->prec != NODE_PREC_FP32) {
      const int np = h->planes;
      std::vector<uint16_t> r16(cnt * np), c16(cnt * np);
      auto tobf = [](float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
        u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
        return (uint16_t)(u >> 16);
      };
      auto frombf = [](uint16_t b) {
        uint32_t u = (uint32_t)b << 16;
        float f;
        std::memcpy(&f, &u, 4);
        return f;
      };
      for (size_t k = 0; k < cnt; ++k) {
        r16[k] = tobf(wr[k]);
        c16[k] = tobf(wc[k]);
        if (np == 2) {  // lo plane: the rounding residual
          r16[cnt + k] = tobf(wr[k] - frombf(r16[k]));
          c16[cnt + k] = tobf(wc[k] - frombf(c16[k]));
        }
      }
      HIPCHK(h, hipMemcpyAsync(ly.Wr, r16.data(), cnt * 2 * np, hipMemcpyHostToDevice, h->stream));
      HIPCHK(h, hipMemcpyAsync(ly.Wc, c16.data(), cnt * 2 * np, hipMemcpyHostToDevice, h->stream));
      HIPCHK(h, hipMemcpyAsync(ly.bias, b.data(), ly.outp * 4, hipMemcpyHostToDevice, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));  // host staging vectors end with this layer
    } else {
      HIPCHK(h, hipMemcpyAsync(ly.Wr, wr.data(), cnt * 4, hipMemcpyHostToDevice, h->stream));
      HIPCHK(h, hipMemcpyAsync(ly.Wc, wc.data(), cnt * 4, hipMemcpyHostToDevice, h->stream));
      HIPCHK(h, hipMemcpyAsync(ly.bias, b.data(), ly.outp * 4, hipMemcpyHostToDevice, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
    }
  }
  if (!h->pflat && dalloc_t(h, &h->pflat, h->P * sizeof(float))) return NODE_ERR_HIP;
  HIPCHK(h, hipMemcpyAsync(h->pflat, p, h->P * sizeof(float), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->have_params = true;
  h->have_forward = false;
  h->rg.have = false;
  return 0;
}

int32_t node_set_params(node_handle_t* h, const float* p, int64_t n) {
  if (!h) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  return set_params_host(h, p, n);
}

int32_t node_set_params_device(node_handle_t* h, const float* d_p, int64_t n) {
  if (!h) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  if (n != h->P) return fail(h, NODE_ERR_INVALID, "parameter count mismatch");
  std::vector<float> p(n);
  HIPCHK(h, hipMemcpyAsync(p.data(), d_p, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return set_params_host(h, p.data(), n);
}

int32_t node_forward_device(node_handle_t* h, const node_solve_opts* opts, const float* d_u0,
                            const double* saveat, int64_t n_saveat, float* d_out,
                            int64_t out_cap, node_stats* stats) {
  if (!h) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  if (!d_u0) return fail(h, NODE_ERR_INVALID, "u0 is NULL");
  if (h->cnf) return fail(h, NODE_ERR_UNSUPPORTED, "CNF handle: use node_cnf_forward");
  if (n_saveat < 0 || (n_saveat > 0 && !saveat)) return fail(h, NODE_ERR_INVALID, "bad saveat");
  int rc = forward_impl(h, opts, d_u0, saveat, n_saveat, stats);
  if (rc) return rc;
  if (d_out) {
    if (h->nsaved > out_cap) return fail(h, NODE_ERR_CAPACITY, "out_cap smaller than the number of saved states");
    HIPCHK(h, hipMemcpyAsync(d_out, h->out_dev, (size_t)h->nsaved * h->B * h->D * sizeof(float),
                             hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  return 0;
}

int32_t node_forward(node_handle_t* h, const node_solve_opts* opts, const float* u0,
                     const double* saveat, int64_t n_saveat, float* out, int64_t out_cap,
                     node_stats* stats) {
  if (!h) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  if (!u0) return fail(h, NODE_ERR_INVALID, "u0 is NULL");
  if (h->cnf) return fail(h, NODE_ERR_UNSUPPORTED, "CNF handle: use node_cnf_forward");
  if (n_saveat < 0 || (n_saveat > 0 && !saveat)) return fail(h, NODE_ERR_INVALID, "bad saveat");
  if (int rc = maybe_autotune(h)) return rc;  // before u0 is staged in the scratch buffer
  float* du0 = h->scratch;  // u0 staged through the scratch state buffer
  HIPCHK(h, hipMemcpyAsync(du0, u0, (size_t)h->B * h->D * sizeof(float), hipMemcpyHostToDevice, h->stream));
  int rc = forward_impl(h, opts, du0, saveat, n_saveat, stats);
  if (rc) return rc;
  if (out) {
    if (h->nsaved > out_cap) return fail(h, NODE_ERR_CAPACITY, "out_cap smaller than the number of saved states");
    HIPCHK(h, hipMemcpyAsync(out, h->out_dev, (size_t)h->nsaved * h->B * h->D * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  return 0;
}

int32_t node_cnf_forward(node_handle_t* h, const node_solve_opts* opts, const float* x,
                         const float* probe, float* z, float* delta_logp, float* logpx,
                         node_stats* stats) {
  if (!h) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  if (!h->cnf) return fail(h, NODE_ERR_UNSUPPORTED, "not a CNF handle (input_map NODE_MAP_CNF*)");
  const bool need_probe = !h->cnf_exact || h->cnf_reg;  // the exact trace alone needs none
  if (!opts || !x || (need_probe && !probe)) return fail(h, NODE_ERR_INVALID, "null argument");
  if (!h->have_params) return fail(h, NODE_ERR_STATE, "parameters not set");
  h->cnf_fwd = false;
  if (int rc = ensure_adjoint(h)) return rc;  // Lm holds the probe for the data-gradient GEMMs
  if (int rc = maybe_autotune(h)) return rc;
  const int dm = h->cnf_dm, B = h->B;
  if (!h->probe && dalloc_t(h, &h->probe, (size_t)h->Bp * dm * sizeof(float))) return NODE_ERR_HIP;
  if (!h->fbuf && dalloc_t(h, &h->fbuf, (size_t)h->Bp * dm * sizeof(float))) return NODE_ERR_HIP;
  std::vector<float> ph((size_t)B * dm, 0.f);
  if (probe) std::copy(probe, probe + (size_t)B * dm, ph.begin());
  HIPCHK(h, hipMemcpyAsync(h->probe, ph.data(), ph.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
  StateGeom sgm = geom(h);
  sgm.D = dm;
  const Layer& ll = h->layers[h->L - 1];
  const int64_t pl = (int64_t)h->Bp * ll.outp;
  launch_prep_lambda(sgm, h->probe, h->Lm, h->LmT, ll.outp, h->ldT, pl, (int64_t)h->ldT * ll.outp,
                     h->stream);
  if (h->cnf_exact) {
    // unit-probe seeds e_k at the last layer, and the trace accumulators
    const size_t unit = (size_t)h->Bp * h->esz * h->planes * ll.outp;
    if (h->LmS.empty()) {
      h->LmS.assign(dm, nullptr);
      for (int k = 0; k < dm; ++k)
        if (int rc = dalloc(h, &h->LmS[k], unit)) return rc;
      if (int rc = dalloc_t(h, &h->trc, (size_t)h->Bp * sizeof(float))) return rc;
      if (int rc = dalloc_t(h, &h->wacc, (size_t)h->Bp * dm * sizeof(float))) return rc;
    }
    std::vector<float> oh((size_t)B * dm);
    for (int k = 0; k < dm; ++k) {
      for (int b = 0; b < B; ++b)
        for (int c = 0; c < dm; ++c) oh[(size_t)b * dm + c] = c == k ? 1.f : 0.f;
      HIPCHK(h, hipMemcpyAsync(h->fbuf, oh.data(), oh.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
      launch_prep_lambda(sgm, h->fbuf, h->LmS[k], nullptr, ll.outp, 0, pl, 0, h->stream);
      HIPCHK(h, hipStreamSynchronize(h->stream));  // oh is rewritten for the next seed
    }
  }
  // u0 = [x; 0] staged through the scratch state buffer
  std::vector<float> u0((size_t)B * h->D, 0.f);
  for (int b = 0; b < B; ++b)
    for (int c = 0; c < dm; ++c) u0[(size_t)b * h->D + c] = x[(size_t)b * dm + c];
  HIPCHK(h, hipMemcpyAsync(h->scratch, u0.data(), u0.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
  node_solve_opts o = *opts;
  o.save_start = 0; o.save_end = 1; o.save_everystep = 0;
  node_stats st{};
  h->cnf_active = true;
  int rc = forward_impl(h, &o, h->scratch, nullptr, 0, &st);
  h->cnf_active = false;
  if (rc) return rc;
  std::vector<float> fin((size_t)B * h->D);
  if (st.nsaved >= 1) {
    HIPCHK(h, hipMemcpyAsync(fin.data(), h->out_dev, fin.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  const double lc = 0.5 * dm * std::log(2.0 * M_PI);
  for (int b = 0; b < B; ++b) {
    double ss = 0.0;
    for (int c = 0; c < dm; ++c) {
      const float v = fin[(size_t)b * h->D + c];
      if (z) z[(size_t)b * dm + c] = v;
      ss += (double)v * v;
    }
    const float dl = fin[(size_t)b * h->D + dm];
    if (delta_logp) delta_logp[b] = dl;
    if (logpx) logpx[b] = (float)(-0.5 * ss - lc - (double)dl);
  }
  h->cnf_lam.assign(2 * (size_t)B, 0.f);
  if (h->cnf_reg)
    for (int b = 0; b < B; ++b) {
      h->cnf_lam[b] = fin[(size_t)b * h->D + dm + 1];
      h->cnf_lam[(size_t)B + b] = fin[(size_t)b * h->D + dm + 2];
    }
  h->probe_h = ph;
  h->cnf_fwd = st.retcode == NODE_RET_SUCCESS;
  if (stats) *stats = st;
  return 0;
}

int32_t node_cnf_backward(node_handle_t* h, const float* dL_dz, const float* dL_ddlogp,
                          float* dL_dx, double* dL_dp, node_stats* stats) {
  if (!h) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  if (!h->cnf) return fail(h, NODE_ERR_UNSUPPORTED, "not a CNF handle (input_map NODE_MAP_CNF*)");
  if (!dL_dz || !dL_ddlogp) return fail(h, NODE_ERR_INVALID, "null argument");
  if (!h->cnf_fwd || !h->have_forward)
    return fail(h, NODE_ERR_STATE, "no successful CNF forward solve to differentiate");
  const int dm = h->cnf_dm, B = h->B, L = h->L;
  const int64_t n = (int64_t)B * h->D;
  if (n > h->dLdout_cap) {
    dfree(h, h->dLdout_dev);
    h->dLdout_dev = nullptr;
    if (int rc = dalloc_t(h, &h->dLdout_dev, n * sizeof(float))) return rc;
    h->dLdout_cap = n;
  }
  // cotangent of the saved end state [z; Δlogp] (the regulariser rows, when
  // present, get none: their outputs carry no gradient on this path)
  std::vector<float> cot((size_t)n, 0.f);
  for (int b = 0; b < B; ++b) {
    for (int c = 0; c < dm; ++c) cot[(size_t)b * h->D + c] = dL_dz[(size_t)b * dm + c];
    cot[(size_t)b * h->D + dm] = dL_ddlogp[b];
  }
  HIPCHK(h, hipMemcpyAsync(h->dLdout_dev, cot.data(), cot.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
  StateGeom sgm = geom(h);
  sgm.D = dm;
  const Layer& l0 = h->layers[0];
  const Layer& ll = h->layers[L - 1];
  // per tangent chain k (the probe e, or unit probe e_k of the exact trace):
  // t_{-1} = [e_k; 0] row-major (layer 0's tangent input) and transposed
  // (slice k of X0^T, layer 0's weight gradient), and the reverse seed
  // abar_{L-1} = -λ_Δ e_k (λ_Δ is constant over the adjoint solve) with its
  // transposed copy (slice k of Lm^T); the seeds hold for the whole solve
  std::vector<float> tv((size_t)B * dm), sv((size_t)B * dm);
  for (int k = 0; k < h->npr; ++k) {
    for (int b = 0; b < B; ++b)
      for (int c = 0; c < dm; ++c) {
        const float e = h->cnf_exact ? (c == k ? 1.f : 0.f) : h->probe_h[(size_t)b * dm + c];
        tv[(size_t)b * dm + c] = e;
        sv[(size_t)b * dm + c] = -dL_ddlogp[b] * e;
      }
    HIPCHK(h, hipMemcpyAsync(h->probe, tv.data(), tv.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->fbuf, sv.data(), sv.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
    launch_prep_lambda(sgm, h->probe, h->E0s[k], tan_half(h, h->X0T, k), l0.inp, h->ldT,
                       (int64_t)h->Bp * l0.inp, (int64_t)h->ldT * h->Dp, h->stream);
    launch_prep_lambda(sgm, h->fbuf, h->AbLs[k], tan_half(h, h->LmT, k), ll.outp, h->ldT,
                       (int64_t)h->Bp * ll.outp, (int64_t)h->ldT * h->Dp, h->stream);
    HIPCHK(h, hipStreamSynchronize(h->stream));  // tv / sv are rewritten for the next chain
  }
  // the device probe back to the forward's (the exact trace's regulariser row)
  HIPCHK(h, hipMemcpyAsync(h->probe, h->probe_h.data(), h->probe_h.size() * sizeof(float),
                           hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipGetLastError());
  h->cnf_adj = true;
  int rc = backward_impl(h, h->dLdout_dev, stats);
  h->cnf_adj = false;
  if (rc) return rc;
  std::vector<float> lam0((size_t)n);
  if (int rc2 = copy_grads(h, lam0.data(), dL_dp, hipMemcpyDeviceToHost)) return rc2;
  if (dL_dx)
    for (int b = 0; b < B; ++b)
      for (int c = 0; c < dm; ++c) dL_dx[(size_t)b * dm + c] = lam0[(size_t)b * h->D + c];
  return 0;
}

int32_t node_cnf_regularizers(node_handle_t* h, float* lam1, float* lam2) {
  if (!h) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  if (!h->cnf || !h->cnf_reg)
    return fail(h, NODE_ERR_UNSUPPORTED, "not a CNF handle with NODE_MAP_CNF_REGULARIZE");
  if (!h->cnf_fwd) return fail(h, NODE_ERR_STATE, "no successful CNF forward solve");
  if (lam1) std::copy(h->cnf_lam.begin(), h->cnf_lam.begin() + h->B, lam1);
  if (lam2) std::copy(h->cnf_lam.begin() + h->B, h->cnf_lam.end(), lam2);
  return 0;
}

int32_t node_saved_times(node_handle_t* h, double* t_out, int64_t cap) {
  if (!h || !t_out) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  if (!h->have_forward) return fail(h, NODE_ERR_STATE, "no forward solve");
  if (cap < (int64_t)h->h_save.size()) return fail(h, NODE_ERR_CAPACITY, "cap too small");
  std::copy(h->h_save.begin(), h->h_save.end(), t_out);
  return 0;
}

int32_t node_copy_saved(node_handle_t* h, float* out, int64_t cap) {
  if (!h || !out) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  if (!h->have_forward) return fail(h, NODE_ERR_STATE, "no forward solve");
  if (cap < h->nsaved) return fail(h, NODE_ERR_CAPACITY, "cap too small");
  HIPCHK(h, hipMemcpyAsync(out, h->out_dev, (size_t)h->nsaved * h->B * h->D * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int32_t node_backward_device(node_handle_t* h, const float* d_dLdout, float* d_dLdu0,
                             double* d_dLdp, node_stats* stats) {
  if (!h) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  if (!d_dLdout) return fail(h, NODE_ERR_INVALID, "dLdout is NULL");
  int rc = backward_impl(h, d_dLdout, stats);
  if (rc) return rc;
  return copy_grads(h, d_dLdu0, d_dLdp, hipMemcpyDeviceToDevice);
}

int32_t node_backward(node_handle_t* h, const float* dLdout, float* dLdu0, double* dLdp,
                      node_stats* stats) {
  if (!h) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  if (!dLdout) return fail(h, NODE_ERR_INVALID, "dLdout is NULL");
  if (!h->have_forward) return fail(h, NODE_ERR_STATE, "no forward solve to differentiate");
  const int64_t n = std::max<int64_t>(h->nsaved, 1) * h->B * h->D;
  if (n > h->dLdout_cap) {
    dfree(h, h->dLdout_dev);
    h->dLdout_dev = nullptr;
    if (int rc = dalloc_t(h, &h->dLdout_dev, n * sizeof(float))) return rc;
    h->dLdout_cap = n;
  }
  HIPCHK(h, hipMemcpyAsync(h->dLdout_dev, dLdout, (size_t)h->nsaved * h->B * h->D * sizeof(float),
                           hipMemcpyHostToDevice, h->stream));
  int rc = backward_impl(h, h->dLdout_dev, stats);
  if (rc) return rc;
  return copy_grads(h, dLdu0, dLdp, hipMemcpyDeviceToHost);
}

int32_t node_ragged_forward(node_handle_t* h, const node_solve_opts* opts, int64_t n,
                            const double* t1, const float* u0, const int64_t* save_offsets,
                            const double* save_times, float* out, int32_t* retcodes,
                            int32_t* traj_stats, node_ragged_stats* stats) {
  if (!h) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  return ragged_forward_impl(h, opts, n, t1, u0, save_offsets, save_times, out, retcodes,
                             traj_stats, stats);
}

int32_t node_ragged_backward(node_handle_t* h, const float* dLdout, float* dLdu0, double* dLdp,
                             float* dLdp_traj, int32_t* retcodes, int32_t* traj_stats,
                             node_ragged_stats* stats) {
  if (!h) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  return ragged_backward_impl(h, dLdout, dLdu0, dLdp, dLdp_traj, retcodes, traj_stats, stats);
}

int32_t node_rhs(node_handle_t* h, const float* u, float* du) {
  if (!h || !u || !du) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  if (h->cnf) return fail(h, NODE_ERR_UNSUPPORTED, "CNF handle: use node_cnf_forward");
  if (!h->have_params) return fail(h, NODE_ERR_STATE, "parameters not set");
  if (int rc = maybe_autotune(h)) return rc;
  const StateGeom sg = geom(h);
  HIPCHK(h, hipMemcpyAsync(h->ucur, u, (size_t)h->B * h->D * sizeof(float), hipMemcpyHostToDevice, h->stream));
  const int64_t px = (int64_t)h->Bp * h->Dp;
  launch_prep_plain(sg, h->ucur, h->X0, nullptr, nullptr, h->Dp, h->ldT, px, px, h->stream);
  rhs_eval(h, 0, SM_PLAIN, GATE_ALWAYS, nullptr);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(du, h->scratch, (size_t)h->B * h->D * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int32_t node_vjp(node_handle_t* h, const float* u, const float* lambda, float* du, double* dp) {
  if (!h || !u || !lambda || !du || !dp) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  if (h->cnf) return fail(h, NODE_ERR_UNSUPPORTED, "CNF handle: use node_cnf_backward");
  if (!h->have_params) return fail(h, NODE_ERR_STATE, "parameters not set");
  if (int rc = ensure_adjoint(h)) return rc;
  if (int rc = maybe_autotune(h)) return rc;
  const StateGeom sg = geom(h);
  const size_t sb = (size_t)h->B * h->D * sizeof(float);
  HIPCHK(h, hipMemcpyAsync(h->ucur, u, sb, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->lam, lambda, sb, hipMemcpyHostToDevice, h->stream));
  const int64_t px = (int64_t)h->Bp * h->Dp;
  const int64_t pxt = (int64_t)h->ldT * h->Dp;
  launch_prep_plain(sg, h->ucur, h->X0, h->X0T, h->ucur, h->Dp, h->ldT, px, pxt, h->stream);
  launch_prep_lambda(sg, h->lam, h->Lm, h->LmT, h->Dp, h->ldT, px, pxt, h->stream);
  if (!h->dp32 && dalloc_t(h, &h->dp32, h->P * sizeof(float))) return NODE_ERR_HIP;
  vjp_eval(h, 0, SM_PLAIN, GATE_ALWAYS, WG_STORE_PTR, 0, 0, 0, 0, nullptr, h->scratch, h->dp32);
  HIPCHK(h, hipGetLastError());
  std::vector<float> p32(h->P);
  HIPCHK(h, hipMemcpyAsync(du, h->scratch, sb, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(p32.data(), h->dp32, h->P * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int64_t i = 0; i < h->P; ++i) dp[i] = -(double)p32[i];  // epilogue stores -g
  return 0;
}

int32_t node_set_option(node_handle_t* h, int32_t key, int64_t value) {
  if (!h) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  if (key == NODE_OPT_GRAPHS) {
    h->use_graphs = value != 0;
  } else if (key == NODE_OPT_AUTOTUNE) {
    h->autotune = value != 0;
    h->tuned_data = false;
    if (!h->autotune) {  // back to the heuristic plans
      h->tuned = false;
      for (auto& row : h->tcfg)
        for (int& c : row) c = -1;
      for (auto& row : h->tks)
        for (int& c : row) c = 1;
      for (Layer& ly : h->layers) {
        GemmPlan p = plan(ly.inp, ly.outp);
        ly.wg_cfg = p.cfg; ly.wg_rt = p.rt; ly.wg_ct = p.ct; ly.wg_ks = ly.wg_ks0;
      }
    } else {
      h->tuned = false;
    }
  } else if (key == NODE_OPT_WGRAD_PRECISION) {
    // BF16X3 operands hold the BF16 rounding in their first plane, so the
    // weight-gradient products can read that plane alone
    const bool ok = value == h->prec || (h->prec == NODE_PREC_BF16X3 && value == NODE_PREC_BF16);
    if (!ok) return fail(h, NODE_ERR_INVALID, "weight-gradient precision must equal the handle's, or be BF16 on a BF16X3 handle");
    h->wg_prec = (int)value;
    h->tuned = false;  // re-tunes the weight-gradient plans only (tuned_data stays)
  } else if (key == NODE_OPT_SENSEALG) {
    if (value < NODE_SENSE_INTERPOLATING || value > NODE_SENSE_BACKSOLVE_NOCKPT)
      return fail(h, NODE_ERR_INVALID, "unknown sensealg (NODE_SENSE_*)");
    if (h->cnf && value != NODE_SENSE_INTERPOLATING)
      return fail(h, NODE_ERR_UNSUPPORTED, "sensealg: a CNF handle supports InterpolatingAdjoint only");
    if (value == h->sense) return 0;  // unchanged: the captured graphs stay valid
    h->sense = (int)value;
  } else if (key == NODE_OPT_GRAPH_ATTEMPTS) {
    if (value < 1 || value > 16) return fail(h, NODE_ERR_INVALID, "graph attempts must be 1..16");
    h->graph_attempts = (int)value;
  } else if (key == NODE_OPT_CHECKPOINTING) {
    if (h->cnf && value != 0)
      return fail(h, NODE_ERR_UNSUPPORTED, "checkpointing: not available on a CNF handle");
    h->ckpt = value != 0;
    return 0;  // read by the next forward; the captured graphs stay valid
  } else if (key == NODE_OPT_TAPE_BUDGET) {
    if (value < 0) return fail(h, NODE_ERR_INVALID, "tape budget must be >= 0 bytes");
    h->tape_budget = value;
    return 0;
  } else if (key == NODE_OPT_RAGGED_VJP_PRECISION) {
    if (value != NODE_PREC_FP32 && value != NODE_PREC_BF16)
      return fail(h, NODE_ERR_INVALID, "per-trajectory VJP precision must be NODE_PREC_FP32 or NODE_PREC_BF16");
    h->rg_vjp_prec = (int)value;
    return 0;  // read by the next node_ragged_backward
  } else {
    return fail(h, NODE_ERR_INVALID, "unknown option");
  }
  if (h->fwd_graph) { (void)hipGraphExecDestroy(h->fwd_graph); h->fwd_graph = nullptr; }
  if (h->adj_graph) { (void)hipGraphExecDestroy(h->adj_graph); h->adj_graph = nullptr; }
  return 0;
}

int32_t node_get_plan(node_handle_t* h, int32_t role, int32_t layer, int32_t* tile_rows,
                      int32_t* tile_cols, int32_t* ksplit) {
  if (!h || role < 0 || role > 5 || layer < 0 || layer >= h->L) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  int cfg, ks = 1;
  if (role == PR_WGRAD) {
    cfg = h->layers[layer].wg_cfg;
    ks = h->layers[layer].wg_ks;
  } else {
    const Layer& ly = h->layers[layer];
    int rows = h->Bp, cols = ly.outp, minc = 0;
    if (role == PR_BWD_DATA) cols = ly.inp;
    if (role == PR_FWD_STAGE || role == PR_BWD_LAST) { minc = 1; cols = role == PR_BWD_LAST ? h->Dp : ly.outp; }
    cfg = plan_role(h, role, layer, rows, cols, minc).cfg;
    if (role == PR_FWD_STAGE || role == PR_BWD_LAST) ks = h->tks[role][layer];
  }
  if (tile_rows) *tile_rows = gemm_cfg_tile_rows(cfg);
  if (tile_cols) *tile_cols = gemm_cfg_tile_cols(cfg);
  if (ksplit) *ksplit = ks;
  return 0;
}

int32_t node_rhs_device_noalloc(node_handle_t* h, const float* d_u, float* d_du, int64_t n) {
  if (!h || !d_u || !d_du || n < 1) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  if (h->cnf) return fail(h, NODE_ERR_UNSUPPORTED, "CNF handle: use node_cnf_forward");
  if (!h->have_params) return fail(h, NODE_ERR_STATE, "parameters not set");
  if (!h->na_scratch) return fail(h, NODE_ERR_STATE, "no-alloc array address not resolved");
  // layout in the array: X0 [planes][Bp][Dp], two hidden buffers
  // [planes][Bp][max outp], the FP32 output [Bp][D]; 256-byte aligned parts
  const int64_t Bp = rup(n, BPAD);
  int64_t hmax = 0;
  for (int l = 0; l + 1 < h->L; ++l) hmax = std::max<int64_t>(hmax, h->layers[l].outp);
  int64_t off = 0;
  auto take = [&off](int64_t bytes) { const int64_t o = off; off = rup(off + bytes, 256); return o; };
  const int64_t ox = take((int64_t)h->planes * Bp * h->Dp * h->esz);
  const int64_t oh0 = take((int64_t)h->planes * Bp * hmax * h->esz);
  const int64_t oh1 = take((int64_t)h->planes * Bp * hmax * h->esz);
  const int64_t zero_bytes = off;
  const int64_t oo = take(Bp * h->D * (int64_t)sizeof(float));
  if (off > (int64_t)NOALLOC_BYTES)
    return fail(h, NODE_ERR_CAPACITY,
                "node_rhs_device_noalloc: " + std::to_string(n) + " trajectories need " +
                    std::to_string(off) + " bytes of the " + std::to_string(NOALLOC_BYTES) +
                    "-byte no-alloc array");
  char* base = (char*)h->na_scratch;
  // padded feature columns must read as zeros in the GEMM K loops; the array
  // keeps whatever the last call (of any handle) left there
  launch_fill_zero(base, zero_bytes, h->stream);
  StateGeom sg = geom(h);
  sg.B = (int)n;
  sg.Bp = (int)Bp;
  sg.sv = Bp * h->D;
  sg.ent = 8 * sg.sv;
  const int64_t px = Bp * h->Dp;
  launch_prep_plain(sg, d_u, base + ox, nullptr, nullptr, h->Dp, 0, px, 0, h->stream);
  const void* A = base + ox;
  for (int l = 0; l + 1 < h->L; ++l) {
    const Layer& ly = h->layers[l];
    void* Y = base + ((l & 1) ? oh1 : oh0);
    GemmArgs ga{};
    ga.A = A; ga.lda = ly.inp; ga.B = ly.Wr; ga.ldb = ly.inp; ga.K = ly.inp;
    ga.pa = Bp * ly.inp; ga.pb = (int64_t)ly.outp * ly.inp; ga.ksplit = 1;
    const GemmPlan p = plan((int)Bp, ly.outp, 0);
    ga.row_tiles = p.rt; ga.col_tiles = p.ct;
    const EpiHidden e{ly.bias, ly.act, ly.out, ly.outp, 0, Y, nullptr, nullptr,
                      Bp * ly.outp, 0};
    launch_gemm_hidden(h->prec, p.cfg, ga, e, h->ctl, GATE_ALWAYS, h->stream);
    A = Y;
  }
  const Layer& ly = h->layers[h->L - 1];
  GemmArgs ga{};
  ga.A = A; ga.lda = ly.inp; ga.B = ly.Wr; ga.ldb = ly.inp; ga.K = ly.inp;
  ga.pa = Bp * ly.inp; ga.pb = (int64_t)ly.outp * ly.inp; ga.ksplit = 1;
  const GemmPlan p = plan((int)Bp, ly.outp, 1);
  ga.row_tiles = p.rt; ga.col_tiles = p.ct;
  EpiFwdStage e{};
  e.bias = ly.bias; e.act = ly.act; e.D = h->D; e.B = (int)n; e.stage = 0; e.mode = SM_PLAIN;
  e.tape = nullptr; e.ent = sg.ent; e.sv = sg.sv;
  e.X0 = nullptr; e.ldx = h->Dp; e.map = h->desc.input_map;
  e.scratch = (float*)(base + oo); e.part = nullptr; e.px = px;
  launch_gemm_fwdstage(h->prec, p.cfg, ga, e, h->ctl, GATE_ALWAYS, h->stream);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(d_du, base + oo, (size_t)n * h->D * sizeof(float),
                           hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int32_t node_prof_enable(node_handle_t* h, uint64_t key_mask) {
  if (!h) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  h->prof_mask = key_mask;
  return 0;
}

int32_t node_prof_collect(node_handle_t* h, int32_t* keys, float* ms, int64_t cap,
                          int64_t* n_out) {
  if (!h || !n_out) return NODE_ERR_INVALID;
  std::lock_guard<std::mutex> g(h->mtx);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  int64_t n = 0;
  for (auto& r : h->prof) {
    float t = 0.f;
    HIPCHK(h, hipEventElapsedTime(&t, r.a, r.b));
    if (n < cap && keys && ms) { keys[n] = r.key; ms[n] = t; }
    ++n;
    h->ev_pool.push_back(r.a);
    h->ev_pool.push_back(r.b);
  }
  h->prof.clear();
  *n_out = n;
  return 0;
}

}  // extern "C"
