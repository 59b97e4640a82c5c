/*
 * node_hip.h -- C ABI of the MI355X (gfx950) NeuralODE hot path.
 *
 * Replaces, for supported models, what the reference reaches through
 *
 *   (n::NeuralODE)(x, p, st)                      src/neural_de.jl:47-61
 *     -> solve(prob, Tsit5(); sensealg = InterpolatingAdjoint(
 *              autojacvec = ZygoteVJP()), kwargs...)  src/neural_de.jl:55-58
 *
 * i.e. the OrdinaryDiffEq Tsit5 forward solve (tableau, embedded error, PI
 * controller, initial dt, tstops, dense output, saveat) and the
 * SciMLSensitivity InterpolatingAdjoint backward solve with the Lux Dense
 * chain as right-hand side (model built at src/neural_de.jl:48-50).
 *
 * Conventions
 *  - State arrays use Julia's column-major D x B layout (one contiguous
 *    column of D values per trajectory); saved states are D x B x Nt.
 *  - Parameters are the flat Lux/ComponentArray vector: per Dense layer the
 *    out x in column-major weight, then the bias
 *    (test/BasicNeuralDE/neural_de_tests.jl:45).
 *  - Time, step size and controller state are FP64; the returned parameter
 *    gradient is FP64 (the batch reduction of dL/dp is accumulated in FP64).
 *  - Every entry point returns 0 on success or a negative NODE_ERR_* code;
 *    node_last_error() gives the message.  Solver outcomes (MaxIters,
 *    DtLessThanMin, Unstable, DtNaN) are NOT call failures: they come back in
 *    node_stats.retcode with SciMLBase.ReturnCode numbering.
 *  - A handle is single-threaded (one HIP stream per handle); calls are
 *    synchronous with respect to the host.  Distinct handles are independent.
 *  - The library never calls back into the caller.
 */
#ifndef NODE_HIP_H
#define NODE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NODE_HIP_ABI_VERSION 1
#define NODE_HIP_MAX_LAYERS 8

/* activations of a Dense layer (NNlib names) */
enum { NODE_ACT_IDENTITY = 0, NODE_ACT_TANH = 1, NODE_ACT_SOFTPLUS = 2,
       NODE_ACT_RELU = 3, NODE_ACT_SIGMOID = 4 };
/* parameter-free first layer: identity or x -> x.^3 (WrappedFunction,
 * docs/src/examples/neural_ode.md:32) */
enum { NODE_MAP_IDENTITY = 0, NODE_MAP_CUBE = 1,
/* continuous normalizing flow handles (node_cnf_forward): the model maps z
 * (NODE_MAP_CNF, widths[0] == widths[n_layers] == D) or the time-concatenated
 * [z; t] (NODE_MAP_CNF_TIME, widths[0] == D + 1) to dz/dt; the ODE state is
 * [z; log p] (D + 1 <= 64 entries per sample), the last layer is linear. */
       NODE_MAP_CNF = 2, NODE_MAP_CNF_TIME = 3 };
/* flags OR-ed onto NODE_MAP_CNF / NODE_MAP_CNF_TIME (the layer state of
 * src/ffjord.jl:138-178, regularize and monte_carlo):
 *  NODE_MAP_CNF_REGULARIZE : regularize = true (src/ffjord.jl:131-132,
 *    149-169): the ODE state is [z; log p; r1; r2] with r1' = |f|^2 and
 *    r2' = |e^T df/dz| per sample (D + 3 <= 64); node_cnf_regularizers gives
 *    r1(T), r2(T) (the layer's lambda_1, lambda_2; the reference selects them
 *    with selectdim(pred, N, ...), the batch dimension, which for a matrix picks
 *    two sample columns -- the rows are what the regularisers are).  The
 *    adjoint takes no cotangents for them (zero: they carry no gradient).
 *  NODE_MAP_CNF_EXACT_TRACE : monte_carlo = false (src/ffjord.jl:116-126): the
 *    trace of df/dz exactly, from D unit-probe passes per right-hand side
 *    (and D tangent chains per adjoint VJP); the probe argument of
 *    node_cnf_forward is then only read for the r2 regulariser (e^T J, the
 *    reference's randn! draw) and may be NULL without regularize. */
enum { NODE_MAP_CNF_REGULARIZE = 0x10, NODE_MAP_CNF_EXACT_TRACE = 0x20 };
/* GEMM precision of the right-hand side and its VJPs.  State, stage sums and
 * adjoint state stay FP32 in all modes.
 *  FP32   : FP32 operands on FP32 MFMA.
 *  BF16   : BF16 operands on BF16 MFMA, FP32 accumulation.  Activation
 *           rounding makes the RHS jitter at the 2^-9 level, which Tsit5's error
 *           estimate sees at tight tolerances (more steps), and weight rounding
 *           bounds the gradient accuracy to ~1e-3..1e-2.
 *  BF16X3 : every operand split into two BF16 planes (hi + lo); each GEMM runs
 *           three BF16 MFMA passes (hi*hi + lo*hi + hi*lo), keeping ~16
 *           mantissa bits: FP32-like step counts and gradients. */
enum { NODE_PREC_FP32 = 0, NODE_PREC_BF16 = 1, NODE_PREC_BF16X3 = 2 };

/* SciMLBase.ReturnCode numbering */
enum { NODE_RET_DEFAULT = 0, NODE_RET_SUCCESS = 1, NODE_RET_TERMINATED = 2,
       NODE_RET_DTNAN = 3, NODE_RET_MAXITERS = 4, NODE_RET_DTLESSTHANMIN = 5,
       NODE_RET_UNSTABLE = 6 };

/* call failures */
enum { NODE_OK = 0, NODE_ERR_INVALID = -1, NODE_ERR_HIP = -2,
       NODE_ERR_STATE = -3, NODE_ERR_CAPACITY = -4, NODE_ERR_UNSUPPORTED = -5 };

typedef struct node_handle node_handle_t;

/* Model + batch description.  Mirrors a Lux Chain:
 *   Chain([x -> x.^3,] Dense(w0 => w1, act0), ..., Dense(w_{L-1} => w_L, act_{L-1}))
 * with widths[0] == widths[n_layers] == D (the ODE state dimension). */
typedef struct node_desc {
    int32_t n_layers;                          /* number of Dense layers, 1..8 */
    int32_t widths[NODE_HIP_MAX_LAYERS + 1];
    int32_t acts[NODE_HIP_MAX_LAYERS];
    int32_t input_map;                         /* NODE_MAP_* */
    int32_t precision;                         /* NODE_PREC_* */
    int64_t batch;                             /* B: trajectories (columns) */
} node_desc;

/* Solver options: the kwargs of solve(prob, Tsit5(); ...) that this path
 * honours.  Saving follows OrdinaryDiffEq's internal form: `saveat` holds the
 * interior save times (strictly between t0 and t1, increasing), plus the
 * save_start / save_end / save_everystep flags. */
typedef struct node_solve_opts {
    double t0, t1;            /* tspan */
    double abstol, reltol;    /* 0 -> 1e-6 / 1e-3 (OrdinaryDiffEq defaults) */
    int32_t adaptive;         /* 1: PI-controlled steps; 0: fixed dt */
    int32_t save_start, save_end, save_everystep;
    double dt;                /* adaptive: initial dt (0 -> heuristic); fixed: the step */
    double dtmax;             /* 0 -> |t1 - t0| */
    double dtmin;             /* 0 -> eps of the span */
    int64_t maxiters;         /* 0 -> 100000 */
} node_solve_opts;

typedef struct node_stats {
    int64_t nf;               /* batched RHS (forward) or VJP (adjoint) evaluations */
    int64_t naccept, nreject; /* steps */
    int64_t nsaved;           /* saved states written */
    int32_t retcode;          /* NODE_RET_* */
    int32_t reserved;
    double t_final;
} node_stats;

int32_t node_abi_version(void);

/* Number of flat parameters P for a description (no handle needed). */
int64_t node_num_params(const node_desc* desc);

int32_t node_create(const node_desc* desc, node_handle_t** out);
int32_t node_destroy(node_handle_t* h);
const char* node_last_error(const node_handle_t* h);

/* Parameters in flat Lux order (P FP32 values). Host or device pointer. */
int32_t node_set_params(node_handle_t* h, const float* p, int64_t n);
int32_t node_set_params_device(node_handle_t* h, const float* d_p, int64_t n);

/* Forward solve.  u0: D x B.  out: D x B x n_out with capacity out_cap states
 * (n_out is returned in stats->nsaved; with save_everystep it is only known
 * after the solve: NODE_ERR_CAPACITY is returned when out_cap is too small and
 * node_copy_saved() can then fetch everything).  The forward tape needed by
 * node_backward stays in the handle.
 * Replaces: the forward half of src/neural_de.jl:55-58. */
int32_t node_forward(node_handle_t* h, const node_solve_opts* opts,
                     const float* u0, const double* saveat, int64_t n_saveat,
                     float* out, int64_t out_cap, node_stats* stats);
/* Same with u0/out/saveat in device memory (HIP pointers); no host copies. */
int32_t node_forward_device(node_handle_t* h, const node_solve_opts* opts,
                            const float* d_u0, const double* saveat, int64_t n_saveat,
                            float* d_out, int64_t out_cap, node_stats* stats);

/* FFJORD forward (src/ffjord.jl:81-178, __ffjord + __forward_ffjord with
 * monte_carlo = true): solves d[z; log p]/dt = [f(z[, t]); -e^T (df/dz) e]
 * from [x; 0] over opts->t0..t1 with Tsit5 (error norm over all (D+1) x B
 * entries, one step size for the batch), the Hutchinson probe e (D x B) held
 * fixed for the whole solve (the reference redraws randn! on every call:
 * here the caller draws it, e.g. Rademacher, once per solve).  Writes z(T)
 * (D x B), delta_logp = log p(T) (B) and logpx = logpz(z(T)) - delta_logp
 * with the standard-normal base density (B).  Any output may be NULL.
 * Handle: input_map NODE_MAP_CNF or NODE_MAP_CNF_TIME.  Host pointers. */
int32_t node_cnf_forward(node_handle_t* h, const node_solve_opts* opts, const float* x,
                         const float* probe, float* z, float* delta_logp, float* logpx,
                         node_stats* stats);

/* FFJORD backward: the InterpolatingAdjoint (the layer's default sensealg,
 * src/ffjord.jl:143, differentiated through the Zygote pullback of the
 * Hutchinson term, :95-115) of the last node_cnf_forward solve on this
 * handle, with the same probe.  Cotangents dL_dz (D x B) and dL_ddlogp (B)
 * of z(T) and delta_logp; for a loss on logpx, dL/dz = -dL/dlogpx * z and
 * dL/ddelta_logp = -dL/dlogpx.  The adjoint state [lambda_z; lambda_logp; mu]
 * is solved back from T with Tsit5 at the forward's tolerances (error norm
 * over all (D + 1) x B + P entries), the model's VJP including the
 * second-order VJP of e^T (df/dz) e (tangent chain t_l = act'(a_l) (t_{l-1}
 * W_l), its reverse, and the source abar_t . adot . act''(a) on the primal
 * chain); the weight gradients run as one product of K = 2B over the primal
 * and tangent halves.  Writes dL_dx (D x B) and dL_dp (P, FP64); either may
 * be NULL.  Host pointers. */
int32_t node_cnf_backward(node_handle_t* h, const float* dL_dz, const float* dL_ddlogp,
                          float* dL_dx, double* dL_dp, node_stats* stats);

/* r1(T) and r2(T) per sample (B each) of the last node_cnf_forward on a
 * NODE_MAP_CNF_REGULARIZE handle: the integrated |f|^2 and |e^T df/dz| -- the
 * lambda_1, lambda_2 outputs of (n::FFJORD)(x, ps, st) with regularize = true
 * (src/ffjord.jl:161-166).  Either may be NULL. */
int32_t node_cnf_regularizers(node_handle_t* h, float* lam1, float* lam2);

/* Times of the states saved by the last forward (n = stats->nsaved). */
int32_t node_saved_times(node_handle_t* h, double* t_out, int64_t cap);
/* Copy the saved states of the last forward (D x B x n) to host memory. */
int32_t node_copy_saved(node_handle_t* h, float* out, int64_t cap);

/* InterpolatingAdjoint backward solve of the last forward.
 * dLdout: cotangent of the saved states (D x B x nsaved).  Writes dL/du0
 * (D x B, FP32) and dL/dp (P, FP64).
 * Replaces: the pullback SciMLSensitivity attaches to the solve at
 * src/neural_de.jl:55-58 with sensealg InterpolatingAdjoint(ZygoteVJP()). */
int32_t node_backward(node_handle_t* h, const float* dLdout, float* dLdu0,
                      double* dLdp, node_stats* stats);
int32_t node_backward_device(node_handle_t* h, const float* d_dLdout,
                             float* d_dLdu0, double* d_dLdp, node_stats* stats);

/* One batched right-hand-side evaluation du = model(u, p) (D x B), and one
 * VJP (lambda^T df/du, lambda^T df/dp) -- the Lux Chain and its Zygote
 * pullback (src/neural_de.jl:50, ZygoteVJP at :57).  Host pointers. */
int32_t node_rhs(node_handle_t* h, const float* u, float* du);
int32_t node_vjp(node_handle_t* h, const float* u, const float* lambda,
                 float* du, double* dp);

/* One right-hand side du = model(u, p) of n trajectories (d_u, d_du: D x n,
 * device pointers) that allocates nothing: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * result as node_rhs for those trajectories, any n (not only the handle's
 * batch) whose model input and hidden activations fit the library's fixed
 * no-alloc array: 64 MiB per GPU, held until the process ends (up to ~7,700
 * trajectories at D = 64 with 1024-wide split-BF16 hidden layers); beyond it
 * NODE_ERR_CAPACITY.
 * Every call on a GPU shares that GPU's array and clears what it uses:
 * calls that overlap on one GPU (from any handle) race.  Synchronous. */
int32_t node_rhs_device_noalloc(node_handle_t* h, const float* d_u, float* d_du, int64_t n);

/* ---- Per-trajectory ("ragged") solves ----------------------------------
 * n independent trajectories of the handle's model, each its own ODE problem
 * -- what n separate NeuralODE solves of batch 1 (src/neural_de.jl:47-61, one
 * call per trajectory) compute: own tspan end t1[i], own save times, own error
 * norm (RMS over its own D entries), own initial dt and PI-controlled steps --
 * in one persistent launch (no counterpart in the reference, which has no
 * batched form of per-trajectory solves).
 *  opts     : t0 (shared start), tolerances, adaptive / dt / dtmax / dtmin /
 *             maxiters as for node_forward; t1 and the save flags are ignored.
 *  t1       : [n] tspan ends (> t0).
 *  u0       : D x n.
 *  save_offsets : [n + 1], save_offsets[0] = 0; trajectory i saves at
 *             save_times[save_offsets[i] .. save_offsets[i+1]) (increasing,
 *             within [t0, t1[i]]; t0 and t1[i] are saved only when listed, the
 *             saveat-vector rule of OrdinaryDiffEq).
 *  out      : D x save_offsets[n] saved states.
 *  retcodes : [n] NODE_RET_* per trajectory.  traj_stats (optional): [n][4]
 *             RHS evaluations, accepted steps, rejected steps, tape records.
 * Host pointers.  Computes in FP32 (MFMA 16x16x4 FP32) whatever the handle's
 * precision: at the tolerances these solves are run with, BF16 operands push
 * the error estimate onto their rounding floor (the adjoint's VJPs can opt in
 * to BF16 operands: NODE_OPT_RAGGED_VJP_PRECISION).  Model limits: D <= 32, hidden
 * widths <= 128, at most 4 Dense layers. */
typedef struct node_ragged_stats {
    int64_t nf, naccept, nreject;   /* summed over trajectories */
    int64_t records;                /* forward tape records (accepted steps + 1 each) */
    int64_t n_success;              /* trajectories with retcode Success */
} node_ragged_stats;
int32_t node_ragged_forward(node_handle_t* h, const node_solve_opts* opts, int64_t n,
                            const double* t1, const float* u0, const int64_t* save_offsets,
                            const double* save_times, float* out, int32_t* retcodes,
                            int32_t* traj_stats, node_ragged_stats* stats);

/* InterpolatingAdjoint of every trajectory of the last node_ragged_forward
 * (each its own adjoint problem [lambda; mu] with its own error norm over its
 * D + P entries, steps landing on its save times with the cotangent jumps),
 * in one persistent launch.  dLdout: D x save_offsets[n] cotangents of the
 * saved states.  Writes dL/du0 (D x n), dL/dp summed over trajectories (P,
 * FP64, summed in a fixed order: 64 consecutive ranges of trajectories each in
 * trajectory order, then the ranges in order) and optionally each trajectory's own
 * dL/dp (P x n, FP32, the adjoint state mu of that trajectory).  All forward
 * retcodes must be Success.  Host pointers.
 * Replaces: per trajectory, the pullback of src/neural_de.jl:55-58 with
 * sensealg InterpolatingAdjoint(ZygoteVJP()). */
int32_t node_ragged_backward(node_handle_t* h, const float* dLdout, float* dLdu0, double* dLdp,
                             float* dLdp_traj, int32_t* retcodes, int32_t* traj_stats,
                             node_ragged_stats* stats);

/* Execution options (no counterpart in the reference):
 *   NODE_OPT_GRAPHS          1 (default): step attempts are captured once into a
 *                            HIP graph and replayed; 0: launched one by one.
 *   NODE_OPT_GRAPH_ATTEMPTS  step attempts per captured graph (default 4, 1..16). */
/*   NODE_OPT_AUTOTUNE        1 (default): before its first solve a handle times every
 *                            valid tile configuration (and split-K factor of the
 *                            weight gradients) of each GEMM role on its own
 *                            shapes and keeps the fastest; 0: fixed heuristic. */
/*   NODE_OPT_WGRAD_PRECISION NODE_PREC_* of the adjoint's weight-gradient products
 *                            (lambda^T df/dp, a third of the adjoint's GEMM work).
 *                            Default: the handle's precision.  A BF16X3 handle
 *                            accepts NODE_PREC_BF16: one MFMA pass instead of three
 *                            on those products; the data path (state, adjoint state,
 *                            and so the step sizes) keeps BF16X3. */
/*   NODE_OPT_SENSEALG        adjoint of node_backward / node_backward_device
 *                            (the reference's sensealg keyword, src/neural_de.jl:57):
 *                            NODE_SENSE_INTERPOLATING (default) InterpolatingAdjoint,
 *                            u(t) from the forward tape; NODE_SENSE_BACKSOLVE
 *                            BacksolveAdjoint(checkpointing = true): u integrated
 *                            back from u(T) with the adjoint, reset to the forward's
 *                            saved state at each saved time; NODE_SENSE_BACKSOLVE_NOCKPT
 *                            BacksolveAdjoint(checkpointing = false). */
/*   NODE_OPT_CHECKPOINTING   1: InterpolatingAdjoint(checkpointing = true) (the
 *                            reference re-exports it, src/DiffEqFlux.jl:53; the
 *                            sensealg keyword at src/neural_de.jl:57).  The next
 *                            node_forward keeps no step record, only the checkpoint
 *                            states (u0, the saved states, u(t1)) and a ring of four
 *                            tape entries; node_backward re-solves the forward over
 *                            each interval between consecutive checkpoints (t0, the
 *                            saved times inside the span, t1), last interval first,
 *                            from its checkpoint state with the forward's tolerances
 *                            (initial dt: the heuristic, then the last step of the
 *                            previous re-solve), and integrates the adjoint over it
 *                            with u(t) from that re-solve.  Device memory then
 *                            depends on the longest interval, not on the whole
 *                            solve.  Not with save_everystep; not on a CNF handle;
 *                            applies to NODE_SENSE_INTERPOLATING only.
 *   NODE_OPT_TAPE_BUDGET     bytes the forward tape may take (0: unlimited, the
 *                            default).  A solve whose tape would exceed it fails with
 *                            NODE_ERR_CAPACITY (with checkpointing, only an interval
 *                            re-solve can need more than a few entries). */
/*   NODE_OPT_RAGGED_VJP_PRECISION  NODE_PREC_FP32 (default) or NODE_PREC_BF16: the
 *                            products of node_ragged_backward's VJPs on BF16 operands
 *                            (one v_mfma_f32_16x16x16_bf16 per 16x16x16 block instead of
 *                            four FP32 ones; FP32 accumulation).  The forward, the
 *                            parameter-gradient sums and all state stay FP32. */
enum { NODE_OPT_GRAPHS = 1, NODE_OPT_GRAPH_ATTEMPTS = 2, NODE_OPT_AUTOTUNE = 3,
       NODE_OPT_WGRAD_PRECISION = 4, NODE_OPT_SENSEALG = 5, NODE_OPT_CHECKPOINTING = 6,
       NODE_OPT_TAPE_BUDGET = 7, NODE_OPT_RAGGED_VJP_PRECISION = 8 };
enum { NODE_SENSE_INTERPOLATING = 0, NODE_SENSE_BACKSOLVE = 1, NODE_SENSE_BACKSOLVE_NOCKPT = 2 };
int32_t node_set_option(node_handle_t* h, int32_t key, int64_t value);
/* Tile plan chosen for a GEMM role (roles as for node_prof_enable below) and
 * layer: tile rows x cols and split-K factor. */
int32_t node_get_plan(node_handle_t* h, int32_t role, int32_t layer, int32_t* tile_rows,
                      int32_t* tile_cols, int32_t* ksplit);

/* Profiling hook (no counterpart in the reference; for measurement tools).
 * Records a HIP-event pair on the handle's stream around every GEMM launch
 * whose key (role * 8 + layer) has its bit set in key_mask; 0 disables.
 * Roles: 0 forward hidden layer, 1 forward last layer (stage epilogue),
 * 2 adjoint forward recompute, 3 adjoint data gradient, 4 weight gradient,
 * 5 adjoint layer-1 data gradient (stage epilogue). */
int32_t node_prof_enable(node_handle_t* h, uint64_t key_mask);
/* Per-launch durations (ms) and keys recorded since enabling; clears them. */
int32_t node_prof_collect(node_handle_t* h, int32_t* keys, float* ms, int64_t cap,
                          int64_t* n_out);

#ifdef __cplusplus
}
#endif
#endif /* NODE_HIP_H */
