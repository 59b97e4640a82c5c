// Internal declarations shared by the device kernels (kernels.hip) and the
// host driver (node_hip.cpp).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nodehip {

constexpr int MAXL = 8;
constexpr int NSTAGE = 7;
constexpr int MAX_PART = 8192;   // partial-sum slots per section

// Gates: a kernel launched in a speculative step returns at once when its
// gate does not hold (the solve is finished, or no FSAL re-evaluation).
enum Gate { GATE_ALWAYS = 0, GATE_ACTIVE = 1, GATE_FSAL = 2 };

// Epilogue modes of the last forward layer (k = model(u))
enum StageMode { SM_STEP = 0, SM_INIT_F0 = 1, SM_INIT_F1 = 2, SM_PLAIN = 3 };

// Device-resident solver control block.  FP64 time / step / controller.
struct Ctl {
  double t, dt, tprev, dtpropose, qold, q11, eest;
  double tdir, dtmax_abs, dtmin, abstol, reltol, t_end;
  double dt0, d0, d1;          // initial-dt heuristic scratch
  double dt_step;              // dt of the last accepted step (saves / mu commit)
  double dt_fixed;
  int64_t n;                   // accepted steps = tape index of the current step
  int64_t iter, naccept, nreject, nf, maxiters;
  int64_t n_save;              // number of save (forward) / callback (adjoint) times
  int64_t save_next;           // forward: next save index; adjoint: next callback (descending)
  int64_t save_lo, save_hi;    // forward saves pending at the next step_begin
  int64_t save_step;           // tape index the pending saves interpolate in
  int64_t n_fwd;               // adjoint: number of forward steps on the tape
  int32_t done, retcode, accepted, fsal_reeval, adaptive, last_accept;
  int32_t jump_idx;            // adjoint: cotangent index to add at step_begin (-1 none)
  int32_t gsel;                // adjoint: current FSAL parameter-gradient buffer
  int32_t commit_mu;           // adjoint: step_begin adds dt_step*acc_b to mu
  int32_t pad_;                // initial-dt heuristic returned early (small dt)
  int32_t has_dt_;             // dt given by the caller: skip the heuristic
  int32_t gsel_mu_;            // adjoint: current mu buffer (0/1)
  // adjoint: forward-interpolation data for each stage time of the next
  // attempt (slots 0..6) and for the initial-dt probe (slot 7)
  int32_t st_step[8];
  int32_t st_exact[8];         // 1: exact stored state tape[st_step].uprev
  float st_dt[8];
  float st_w[8][7];
  // forward: tape entries form a ring of this many slots (0: one entry per step)
  int64_t tape_ring;
  // adjoint with forward-state checkpoints: the solve pauses (done = paused =
  // 1) after an accepted step lands on pause_t, so that the host can re-solve
  // the forward interval below it before the adjoint continues
  double pause_t;
  int32_t has_pause, paused;
};

struct GemmArgs {
  const void* A; int64_t lda;  // rows x K, K contiguous
  const void* B; int64_t ldb;  // cols x K, K contiguous
  int64_t pa, pb;              // plane strides (elements) of the hi/lo planes (split BF16)
  int K;                       // multiple of the K tile (times ksplit)
  int row_tiles, col_tiles;
  int ksplit;                  // split-K factor (weight gradients of narrow layers)
  // split-K with serial fixup (stage epilogues): each split stores its partial
  // tile in ws, the last split of a tile (per-tile counter) sums the partials in
  // split order and runs the epilogue; counters return to zero afterwards
  float* ws;
  int* cnt;
};

// hidden layer forward: Y = act(acc + bias) (+ transposed copy, + act')
struct EpiHidden {
  const float* bias; int act; int C; int ldo; int ldt;
  void* Y; void* YT; void* S;
  int64_t py, pyt;             // plane strides of Y / YT
};

// last forward layer: k_s = acc + bias, then the next Tsit5 stage input
struct EpiFwdStage {
  const float* bias; int act; int D; int B; int stage; int mode;
  float* tape; int64_t ent; int64_t sv;   // floats per tape entry / per state
  void* X0; int ldx; int map;
  float* scratch;                          // initdt f1 / plain rhs output
  double* part;
  int64_t px;                              // plane stride of X0
};

// continuous normalizing flow stage (state [z; log p], D = dm + 1 columns):
// the GEMM computes w = (e^T df/dx) per sample (x the model input); the
// epilogue forms trace = sum_{c<dm} w e, the stage k = [f; -trace] (f from
// the last layer, FP32 [Bp][dm]) and the Tsit5 stage combination as
// EpiFwdStage does, writing the next model input [z_s; t_s].
struct EpiCnfStage {
  int D; int B; int dm; int tcat; int stage; int mode;
  float* tape; int64_t ent; int64_t sv;
  void* X0; int ldx; int64_t px;
  const float* f; const float* e;          // [Bp][dm] FP32 model output / probe
  double* part;
  // regularize: the state is [z; log p; r1; r2] with r1' = |f|^2, r2' = |w|
  // (w = e^T df/dz); exact trace: npass unit-probe passes (pass = this one's
  // probe index), the GEMM giving row `pass` of the Jacobian, accumulated in
  // trc [Bp] (trace) and wacc [Bp][dm] (e^T J); the last pass forms the stage
  int reg, pass, npass;
  float* trc; float* wacc;
};

// hidden data gradient: delta = acc * act'(z)  (row-major + transposed);
// bsum (optional): FP32 column sums of delta per row tile, [row_tiles][ldo]
// (the layer's bias gradient before the batch reduction over row tiles)
// yact >= 0: act' is derived from the layer output Y (planes py apart) instead
// of read from S (activations whose derivative is a function of the output)
struct EpiBwdData {
  int C; int ldo; int ldt; const void* S; void* Dl; void* DlT;
  int64_t pd, pdt;             // plane strides of Dl / DlT
  float* bsum;
  const void* Y; int64_t py; int yact;
};

// FFJORD adjoint, primal data gradient with the Hutchinson term's
// second-order source: delta = acc * act'(z) + src (src FP32 [rows][ldo])
struct EpiBwdSrc {
  EpiBwdData d;
  const float* src;
};

// FFJORD adjoint, tangent-backward chain into hidden layer l (rows = batch):
// tb = acc, ab = tb * act'_l written row-major + transposed (the transposed
// copy is the second half of the K = 2B weight-gradient operand), and
// src = tb * (adot_l act''_l) in FP32 from the stored tangent t_l = act'_l adot_l
// and output y_l (tanh: -2 y t; softplus: t (1 - act'); sigmoid: t (1 - 2 y))
struct EpiTanBwd {
  int C; int ldo; int ldt; int act;
  const void* Y; const void* S; const void* Tf;  // layer output, act' (S type), tangent
  int64_t py, ptf;                               // plane strides of Y / Tf
  void* Ab; void* AbT; int64_t pab, pabt;
  float* src;
  int src_acc;                                   // add to src (later probes of an exact trace)
};

// layer-1 data gradient: k^lambda_s = -(acc * map'(u)), next stage inputs
struct EpiBwdLast {
  int D; int B; int stage; int mode;
  float* kl; int64_t sv;                   // 8 stage slots of B_pad x D floats
  float* lam; float* lam_new; float* Ucur;
  void* L; void* LT; int ldl; int ldlt;    // lambda stage input (row-major + transposed)
  void* X0; void* X0T; int ldx; int ldxt; int map;
  const float* tape; int64_t ent;
  double* part;
  float* plain_out;                        // node_vjp: write -k (= lambda^T df/du)
  int64_t pl, plt, px, pxt;                // plane strides of L / LT / X0 / X0T
  // BacksolveAdjoint (ubs != nullptr; elementwise stage kernel only): fu =
  // f(u_s) of this stage, ku its stage slots, ubs / ubs_new as in AdjBufs
  const float* fu; float* ku; float* ubs; float* ubs_new;
  // FFJORD adjoint (cnf; elementwise stage kernel only): state [z; Δlogp]
  // (D = dm + 1), model input [z; t] when tcat; k of the Δlogp column is 0
  // (its lambda is constant) and the lambda input of the model is z's part
  int cnf; int dm; int tcat;
};

// split-K partial products of a skinny stage GEMM (FP32 [ksplit][rows][ld])
struct EpiPartial {
  float* ws; int ld; int64_t pstride;
};

// weight gradient (K = batch): fold into the mu accumulators
enum WgradFlags { WG_INIT = 1, WG_B = 2, WG_E = 4, WG_STORE_CUR = 8,
                  WG_STORE_NEXT = 16, WG_STORE_PTR = 32 };
struct EpiWgrad {
  int in, out; int64_t w0;
  float* accb; float* acce; float* gf0; float* gf1; float* gptr;
  float cb, ce, cb1, ce1; int flags;
  float* wpart;                // split-K partials [ksplit][in*out] (fold kernel reduces)
};

// fold kernel work: bias rows of every layer + split-K partial sums of the
// narrow layers' weight gradients
struct FoldArgs {
  int nl;                      // layers
  const void* dT[MAXL];        // transposed delta of each layer's output (K = batch)
  int64_t dplane[MAXL];
  int outs[MAXL];
  int64_t b0s[MAXL];
  int bias_rows;
  // per-row-tile column sums of delta from the data-gradient epilogue
  // (nullptr: the bias gradient sums the transposed delta over the batch)
  const float* bpart[MAXL];
  int bpart_rt[MAXL], bpart_ld[MAXL];
  int nbw;                     // wave-per-row bias blocks (set by the launcher)
  int ldt, Bp;
  int nsk;                     // narrow layers with split-K partials
  int64_t sk_w0[MAXL], sk_n[MAXL];  // flat offset / entry count
  float* sk_part[MAXL];
  int sk_split[MAXL];
};

// tableau in constant memory (stream-ordered copy, waited for)
hipError_t upload_tableau(hipStream_t s);

// ---- launchers (kernels.hip) ----
template <typename T> struct TypeTag {};

void launch_gemm_hidden(int prec, int cfg, const GemmArgs& g, const EpiHidden& e,
                        const Ctl* ctl, int gate, hipStream_t s);
void launch_gemm_fwdstage(int prec, int cfg, const GemmArgs& g, const EpiFwdStage& e,
                          const Ctl* ctl, int gate, hipStream_t s);
void launch_gemm_bwddata(int prec, int cfg, const GemmArgs& g, const EpiBwdData& e,
                         const Ctl* ctl, int gate, hipStream_t s);
void launch_gemm_cnfstage(int prec, int cfg, const GemmArgs& g, const EpiCnfStage& e,
                          const Ctl* ctl, int gate, hipStream_t s);
void launch_gemm_bwdlast(int prec, int cfg, const GemmArgs& g, const EpiBwdLast& e,
                         const Ctl* ctl, int gate, hipStream_t s);
void launch_gemm_wgrad(int prec, int cfg, const GemmArgs& g, const EpiWgrad& e,
                       const Ctl* ctl, int gate, hipStream_t s);
void launch_gemm_partial(int prec, int cfg, const GemmArgs& g, const EpiPartial& e,
                         const Ctl* ctl, int gate, hipStream_t s);
void launch_gemm_tanbwd(int prec, int cfg, const GemmArgs& g, const EpiTanBwd& e,
                        const Ctl* ctl, int gate, hipStream_t s);
void launch_gemm_bwdsrc(int prec, int cfg, const GemmArgs& g, const EpiBwdSrc& e,
                        const Ctl* ctl, int gate, hipStream_t s);
// elementwise stage kernels over summed split-K partials (rows x cols, cols a
// multiple of 64, rows a multiple of 16); one partial-sum slot per block
int stage_grid(int rows, int cols);
void launch_fwd_stage_kernel(int prec, const EpiFwdStage& e, const float* ws, int ks,
                             int64_t pstride, int ldw, int rows, int cols, const Ctl* ctl,
                             int gate, hipStream_t s);
void launch_bwd_stage_kernel(int prec, const EpiBwdLast& e, const float* ws, int ks,
                             int64_t pstride, int ldw, int rows, int cols, const Ctl* ctl,
                             int gate, hipStream_t s);
int gemm_cfg_tile_rows(int cfg);
int gemm_cfg_tile_cols(int cfg);
int gemm_cfg_count();

// elementwise / control kernels
struct StateGeom {
  int B, Bp, D, Dp;          // real / padded batch, state dim, padded state dim
  int map, prec;             // prec: 0 fp32, 1 bf16, 2 split bf16 (hi/lo planes)
  int64_t sv;                // floats per state vector (Bp * D)
  int64_t ent;               // floats per tape entry (8 * sv)
  int dm;                    // state columns fed to the model (D; D - 1 for a CNF: not log p)
  int tcat;                  // 1: the model input is [z; t] (column dm of X0 holds t)
  int x0w;                   // row stride of the model-input matrix X0
};

void launch_fwd_init(const StateGeom& sg, const float* u0, float* tape, void* X0,
                     int64_t px, const Ctl* ctl, double* part, int nblk, hipStream_t s);
void launch_initdt_a(Ctl* ctl, const double* p0, int n0, const double* px, int nx,
                     const double* pm, int nm, double n_elems, hipStream_t s);
void launch_initdt_b(Ctl* ctl, const double* px, int nx, const double* pm, int nm,
                     double n_elems, const double* save_times, int is_adjoint,
                     const double* fwd_times, hipStream_t s);
void launch_fwd_initdt_prep(const StateGeom& sg, const float* tape, void* X0, int64_t px,
                            const Ctl* ctl, hipStream_t s);
void launch_fwd_step_begin(const StateGeom& sg, const float* tape, void* X0, int64_t px,
                           const double* save_times, float* out, const Ctl* ctl,
                           hipStream_t s);
void launch_fwd_controller(Ctl* ctl, const double* part, int n_part,
                           const double* save_times, double* fwd_times, double n_elems,
                           hipStream_t s);
int state_blocks(const StateGeom& sg);

// adjoint
struct AdjBufs {
  float* lam; float* lam_new; float* kl; float* U;
  void* L; void* LT; void* X0; void* X0T;
  int ldl, ldlt, ldx, ldxt;
  int64_t pl, plt, px, pxt;  // plane strides
  const float* tape; const double* fwd_times;
  const double* cb_times; const float* dLdout;
  double* mu; float* accb; float* acce; float* gf0; float* gf1;
  int64_t P;
  // BacksolveAdjoint (ubs != nullptr): u integrated with the adjoint; ubs /
  // ubs_new the state at the step start / its candidate, ku 8 stage slots of
  // f(u), fsaved the forward's saved states (checkpoints when ckpt)
  float* ubs; float* ubs_new; float* ku; const float* fsaved; int ckpt;
};
void launch_adj_init(const StateGeom& sg, const AdjBufs& ab, const Ctl* ctl,
                     double* part, hipStream_t s);
void launch_adj_initdt_prep(const StateGeom& sg, const AdjBufs& ab, const Ctl* ctl,
                            hipStream_t s);
void launch_adj_step_begin(const StateGeom& sg, const AdjBufs& ab, const Ctl* ctl,
                           int gate, hipStream_t s);
void launch_mu_norm(int mode, const AdjBufs& ab, const Ctl* ctl, double* part,
                    int nblk, int gate, hipStream_t s);
int mu_blocks(int64_t P);
void launch_fill_zero(void* p, int64_t bytes, hipStream_t s);
void launch_adj_resume(Ctl* ctl, const Ctl* saved, int64_t n_fwd, double pause_t, int has_pause,
                       const double* fwd_times, hipStream_t s);
void launch_adj_controller(Ctl* ctl, const double* part, int n_lam,
                           const double* part_mu, int n_mu, double n_elems,
                           const double* cb_times, const double* fwd_times,
                           hipStream_t s);
void launch_fold(int prec, const FoldArgs& fa, const EpiWgrad& tmpl, const Ctl* ctl, int gate,
                 hipStream_t s);
void launch_set_stage_interp(Ctl* ctl, int slot, double tau, const double* fwd_times,
                             hipStream_t s);
void launch_prep_plain(const StateGeom& sg, const float* u, void* X0, void* X0T,
                       float* Ucur, int ldx, int ldxt, int64_t px, int64_t pxt, hipStream_t s);
void launch_prep_lambda(const StateGeom& sg, const float* lam, void* L, void* LT,
                        int ldl, int ldlt, int64_t pl, int64_t plt, hipStream_t s);
void launch_gather(const StateGeom& sg, const float* src, int64_t src_stride_floats,
                   float* dst, int64_t n, hipStream_t s);

// ---------------------------------------------------------------- ragged batches
constexpr int RG_MAXL = 4;
constexpr int RG_HT = 8;        // hidden width <= 128 (16-feature tiles)
constexpr int RG_MAXD = 32;     // state dimension <= 32
constexpr int RG_SLOTS = 16;    // trajectory slots per wave (MFMA columns)
constexpr int RG_SUM_CHUNKS = 64;  // trajectory ranges of the dL/dp sum
struct RaggedModel {
  int L, map;
  int in[RG_MAXL], out[RG_MAXL], act[RG_MAXL];
  int ti[RG_MAXL], to[RG_MAXL];       // 16-feature tiles of each layer's input / output
  int pitch[RG_MAXL];                 // LDS row pitch (floats) of W_l: 16*ti + 4
  int lw[RG_MAXL], lb[RG_MAXL];       // LDS offsets (floats) of W_l rows and bias
  int64_t w0[RG_MAXL], b0[RG_MAXL];   // flat parameter offsets
  int lds_floats;
  int sh[RG_MAXL], sd[RG_MAXL];       // adjoint stage block: offsets of h_l (layer input) and delta_l
  int sv;                             // floats per stage block
  int64_t P;
  const float* p;                     // flat FP32 parameters (device)
  // parameter-gradient rows while a trajectory integrates: each layer's
  // weights as rows of rw 16x16 tiles (a tile: lane, then 4 values, the
  // MFMA accumulator order, one 16-byte vector per lane), then each layer's
  // bias; Pn floats per row (>= P, a multiple of 64).  A layer's tile rows run
  // along its longer tile dimension: rowout[l] = 1: row it holds tiles
  // (it, ot = 0..rw-1), else row ot holds tiles (it = 0..rw-1, ot);
  // nrow[l] rows, tiles past the layer's shorter dimension are padding (0).
  // A finished trajectory's row is rewritten in the flat parameter order
  // (first P floats).
  int64_t nw0[RG_MAXL], nb0[RG_MAXL];
  int nrow[RG_MAXL], rowout[RG_MAXL];
  int rw;                             // tiles per row: the longest tile dimension, rounded up to 1, 2, 4 or 8
  int ntile;                          // weight tiles of a row, all layers
  int64_t Pn;
  // adjoint VJP: act' of the forward sweep parked in LDS after the weights,
  // act_stride floats per wave (tile t of layer l at act_off[l] + t, 256
  // floats per tile); 0: parked in the stage block in global memory
  int act_off[RG_MAXL];
  int act_stride;
};
struct RaggedFwdArgs {
  RaggedModel m;
  int D, Dp;
  int64_t n;
  const double* t1;                   // [n] tspan ends
  const float* u0;                    // [n][D]
  const int64_t* save_off;            // [n + 1]
  const double* save_t;               // [save_off[n]]
  float* out;                         // [save_off[n]][D]
  int* retcode;                       // [n]
  int* traj_stats;                    // [n][4]: nf, naccept, nreject, records
  int* last_rec;                      // [n] last tape record of each trajectory
  int* queue;                         // next trajectory to start
  double* rec_t;                      // tape pool: start time of each record
  int* rec_prev;                      // previous record of the same trajectory (-1: first)
  float* rec_data;                    // [rec_cap][8][Dp]: u, k1..k7
  int64_t rec_cap;
  int* rec_count;
  double t0, abstol, reltol, dtmax, dtmin, dt_init;
  int64_t maxiters;
  int adaptive;
  unsigned long long* stats;          // nf, naccept, nreject over all trajectories
};
void launch_ragged_fwd(const RaggedFwdArgs& a, int grid, hipStream_t s);
struct RaggedAdjArgs {
  RaggedModel m;
  int D, Dp;
  int64_t n;
  const double* t1;
  const int64_t* save_off;
  const double* save_t;
  const float* dLdout;                // [save_off[n]][D]
  const int* last_rec;
  const double* rec_t;
  const int* rec_prev;
  const float* rec_data;
  float* dLdu0;                       // [n][D]
  float* dLdp_traj;                   // [n][Pn] per-trajectory parameter gradients (zeroed)
  float* mu_tmp;                      // [slots][Pn]
  float* sbuf;                        // [slots + 1][8][sv] stage blocks (the last: sink of idle columns)
  int64_t nslots_all;                 // slots of the launch (16 per wave)
  int* queue;
  // start order: the trajectory at each queue position (-1: none).  A wave
  // runs trajectories in its first spw slots only (spw <= 16: more waves, each
  // with fewer parameter passes, when there are SIMDs to spare); slot s < spw
  // of wave w starts at position w * spw + s, later starts take positions
  // nslots, nslots + 1, ... from the queue counter
  const int* order;                   // [norder]
  int64_t norder, nslots;
  int spw;
  int vjp_bf16;                       // VJP products on BF16 operands (NODE_OPT_RAGGED_VJP_PRECISION)
  int* retcode;                       // [n]
  int* traj_stats;                    // [n][4]
  double t0, abstol, reltol, dtmax, dtmin, dt_init;
  int64_t maxiters;
  int adaptive;
  unsigned long long* stats;
};
void launch_ragged_adj(const RaggedAdjArgs& a, int grid, int waves, hipStream_t s);
// part: [RG_SUM_CHUNKS][P] FP64 scratch (the per-range sums)
void launch_ragged_sum(const float* rows, int64_t n, int64_t P, int64_t stride, double* part,
                       double* out, hipStream_t s);

}  // namespace nodehip
