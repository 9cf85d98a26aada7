/*
 * ptn_hip.h -- C ABI of the MI355X (gfx950) execution policy for Partensor's
 * dense CPD-ALS.  Plain pointers and sizes only; no exceptions cross it.
 *
 * Layouts are the reference's (Tensor.hpp:42,63-64): tensors column-major with
 * mode 0 fastest, factors I_n x R column-major, "packed" factor buffers =
 * factor 0, factor 1, ... back to back.  Host buffers stay owned by the caller;
 * a handle owns all device memory and streams and is used by one host thread
 * at a time (handles are independent of each other).
 */
#ifndef PTN_HIP_H
#define PTN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTN_MAX_ORDER 8
#define PTN_MAX_RANK 64

typedef struct ptn_handle_s *ptn_handle;

typedef enum {
  PTN_OK = 0,
  PTN_ERR_INVALID = 1,      /* bad argument (dims, rank, pointer, option value) */
  PTN_ERR_OOM = 2,          /* device allocation failed */
  PTN_ERR_HIP = 3,          /* HIP runtime error (message in ptn_last_error) */
  PTN_ERR_UNSUPPORTED = 4,  /* valid request this build does not implement */
  PTN_ERR_STATE = 5         /* e.g. no tensor set on the handle */
} ptn_status;

/* Constraint values = partensor::Constraint (Constants.hpp:49-57). */
enum { PTN_UNCONSTRAINED = 0, PTN_NONNEGATIVITY = 1, PTN_ORTHOGONALITY = 2, PTN_SPARSITY = 3, PTN_CONSTANT = 4 };
/* Tensor storage on the device. */
enum { PTN_TENSOR_F64 = 1, PTN_TENSOR_F32 = 2, PTN_TENSOR_BOTH = 3 };
/* MTTKRP schedule: per-mode MTTKRP as cpd() (Cpd.hpp:324) or the dimension tree as
   cpdDimTree() (CpdDimTree.hpp:355-366). */
enum { PTN_MTTKRP_STANDARD = 0, PTN_MTTKRP_DIMTREE = 2 };
/* Arithmetic of the tensor contractions. */
enum { PTN_ARITH_F64 = 0,      /* FP64 tensor, FP64 matrix cores (the plain path)        */
       PTN_ARITH_F32 = 1,      /* FP32 tensor, FP32 matrix cores, FP32 accumulate        */
       PTN_ARITH_BF16X3 = 2,   /* FP32 tensor, bf16x3 split products, FP32 accumulate    */
       PTN_ARITH_F32_F64 = 3 };/* FP32 tensor, FP64 matrix cores, FP64 accumulate        */

/* Options<Tensor> fields the ALS path consumes (PARTENSOR_basic.hpp:283-336) plus
   the execution choices of this policy. */
typedef struct {
  int      max_iter;            /* DefaultValues: 20   */
  double   threshold_error;     /* 1e-3                */
  double   nesterov_delta_1;    /* 1e-2                */
  double   nesterov_delta_2;    /* 1e-2                */
  int      accel_coeff;         /* 3                   */
  int      accel_fail;          /* 0                   */
  int      acceleration;        /* 1                   */
  int      normalization;       /* 1                   */
  int      constraints[PTN_MAX_ORDER];  /* nonnegativity */
  /* policy choices */
  int      mttkrp_form;         /* PTN_MTTKRP_STANDARD */
  int      arith;               /* PTN_ARITH_F64       */
  int      max_inner;           /* Nesterov cap, NesterovMNLS.hpp:187: 50 (<= 255) */
  int      fuse_accel;          /* dimension tree + FP32: fold the acceleration pass into the next left-node pass */
  /* the learned NNLS warm start (no reference counterpart) */
  int      warm_start;          /* 1: start each NNLS from the handle's network where it passes the KKT test */
  double   warm_kkt_tol;        /* a row takes the network's start if its relative KKT residual
                                   max|min(L a, a Zt - Wt)| / max|Wt| is <= this (default 1e-6) */
  int      warm_refine;         /* Nesterov steps run from every proposal before that test (0 .. max_inner,
                                   default 0); they are counted in nnls_iters */
  int      mlp_harvest;         /* 1: append every NNLS (inputs, solution) pair to the handle's training set */
  int      time_passes;         /* 1: HIP events around the two tensor passes of every sweep (t_pass_ms) */
} ptn_options;

/* Status<Tensor> (PARTENSOR_basic.hpp:556-595) minus the factors, which are written
   to a caller buffer. */
typedef struct {
  double frob_sq;      /* Status.frob_tns = ||X||^2 */
  double f_value;      /* ||X - [[A]]||_F           */
  double rel_cost;     /* f_value / sqrt(frob_sq)   */
  int    ao_iter;
  int    accel_coeff;  /* options as mutated by the acceleration bookkeeping */
  int    accel_fail;
  double sweep_seconds;  /* wall time of the outer loop (device-synchronised) */
} ptn_result;

/* One record per outer iteration (the reference logs iter / fvalue / rel, Cpd.hpp:317). */
typedef struct {
  double f_value, rel_cost, f_accel;
  int    accel_tried, accel_accepted;
  int    nnls_iters[PTN_MAX_ORDER];
  int    zero_col_fallback[PTN_MAX_ORDER];
  double lipschitz[PTN_MAX_ORDER];
  double mu[PTN_MAX_ORDER];
  double t_end;        /* seconds from the start of the outer loop to the end of this sweep (device-synchronised) */
  int    warm_rows[PTN_MAX_ORDER];  /* rows whose NNLS started from the network's proposal */
  double t_pass_ms[2]; /* time_passes: device time of the tensor passes launched in this sweep, [0] the
                          left-node pass (X x_2 A_2, or its pair with the line search), [1] the right-node
                          pass (dimension tree, FP32 contractions, order 3); 0 where none ran */
  int    warm_steps[PTN_MAX_ORDER]; /* of nnls_iters, the steps run from the proposals (warm_refine) */
} ptn_sweep_log;

/* ---- handle ---- */
ptn_status  ptn_create(ptn_handle *h, int device);
void        ptn_destroy(ptn_handle h);
const char *ptn_last_error(ptn_handle h);
void        ptn_default_options(ptn_options *o);
int         ptn_abi_version(void);

/* ---- the tensor (uploaded once; replaces the per-block host->managed copies of
        CUDAMTTKRP.hpp:157-163) ---- */
/* X: column-major FP64 host array.  storage: PTN_TENSOR_F64 / _F32 / _BOTH. */
ptn_status ptn_set_tensor(ptn_handle h, int N, const int64_t *dims, const double *X, int storage);
/* Synthetic tensor on the device: X = [[A]] with A drawn from (seed, dist: 0 U[0,1),
   1 Exp(1)) plus Gaussian noise at snr_db (no noise when snr_db is NaN). */
ptn_status ptn_generate_tensor(ptn_handle h, int N, const int64_t *dims, int R_true, uint64_t seed, int dist,
                               double snr_db, int storage);
/* Factors drawn on the device with the same generator, returned packed column-major. */
ptn_status ptn_generate_factors(ptn_handle h, int N, const int64_t *dims, int R, uint64_t seed, int dist,
                                double *packed_out);
ptn_status ptn_tensor_norm_sq(ptn_handle h, double *out);

/* ---- the ALS driver: replaces cpd(execution::cuda, ...) (Cpd.hpp:1438-1467 ->
        CpdCUDA.hpp:277-383) and cpdDimTree(...) (CpdDimTree.hpp:807-825 -> 317-401)
        for a tensor set on the handle and caller-given initial factors
        (cpd(tnsX, R, options, factorsInit), Cpd.hpp:550-577) ---- */
ptn_status ptn_cpd(ptn_handle h, int R, const double *init_packed, const ptn_options *opt,
                   double *out_packed, ptn_result *res, ptn_sweep_log *log, int log_cap);

/* ---- single operations (replace the reference functions named) ---- */
/* mttkrp(), MTTKRP.hpp:60-69 / hybrid_mttkrp, CUDAMTTKRP.hpp:224-338: M (I_mode x R col-major). */
ptn_status ptn_mttkrp(ptn_handle h, int R, const double *factors_packed, int mode, int arith, double *M_out);
/* the N MTTKRPs of one dimension-tree sweep with fixed factors (DimTrees.hpp:668-928), packed. */
ptn_status ptn_mttkrp_dimtree(ptn_handle h, int R, const double *factors_packed, int arith, double *M_packed_out);
/* NesterovMNLS, NesterovMNLS.hpp:181-234: Z (R x R), W, A (I x R col-major); A in/out. */
ptn_status ptn_nesterov_mnls(ptn_handle h, int64_t I, int R, const double *Z, const double *W, double delta_1,
                             double delta_2, int max_inner, double *A, int *iters, double *L, double *mu);
/* No counterpart in the reference: ptn_nesterov_mnls for callers that cannot allocate
   device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  Z (R x R), W and A (I x R, A in/out) are DEVICE pointers, column-major; the
   results equal ptn_nesterov_mnls's bit for bit.  Every other buffer is a static device
   array of PTN_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends, its address
   looked up by ptn_create): a call needing more than that -- about
   (max_inner + 4) * I * 8 * round_up(R, 16) bytes -- fails with PTN_ERR_INVALID.  All
   calls on one GPU share that array: calls that overlap on one GPU (several handles or
   threads) race.  The call synchronises the handle's stream before it returns;
   iters / L / mu may be NULL. */
#define PTN_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
ptn_status ptn_nesterov_mnls_noalloc(ptn_handle h, int64_t I, int R, const double *Z, const double *W,
                                     double delta_1, double delta_2, int max_inner, double *A, int *iters, double *L,
                                     double *mu);
/* ComputeSVD extremes for the symmetric V (NesterovMNLS.hpp:45-52). */
ptn_status ptn_sym_extremes(ptn_handle h, int R, const double *V, double *L, double *mu);
/* Normalize, Normalize.hpp:92-174: grams (N x R x R, col-major each) and factors in/out. */
ptn_status ptn_normalize(ptn_handle h, int N, const int64_t *dims, int R, int weight_factor, double *grams_packed,
                         double *factors_packed);
/* ||X - [[A]]||_F for the handle's tensor (Cpd.hpp:131-135). */
ptn_status ptn_cost_explicit(ptn_handle h, int R, const double *factors_packed, double *f);

/* Average device time (ms) of `iters` back-to-back launches of one contraction
   kernel on the handle's stream, bracketed by HIP events on that stream.
   kernel: 0 = left node X x_{N-1} A_{N-1} (gemm_cm; for an order-4 tensor the fused node
   X_(0,1) (A_3 (.) A_2) with its split-k reduction), 1 = right node X x_0 A_0 (gemm_ck; for an
   order-3 tensor on the FP32-output paths fused with the TTV by A_1, i.e. the mode-2 MTTKRP),
   2 = mode-0 standard MTTKRP.  Operand packing is done before the timed launches.
   bytes_out: algorithmic bytes of one launch (tensor once + output + factor rows). */
ptn_status ptn_time_contraction(ptn_handle h, int kernel, int R, const double *factors_packed, int arith, int iters,
                                double *avg_ms, double *bytes_out);

/* ---- batched independent CPDs.  No counterpart in the reference, which runs one cpd()
        per tensor: many small order-3 non-negative problems of uneven sizes (modes in
        [1, 128]) and ranks ([1, 16]) advanced together, one persistent kernel per ALS
        sweep (the sweep of Cpd.hpp:288-356: standard MTTKRP, NesterovMNLS per factor,
        Gram-form cost, Normalize, line_search_accel), tensors in FP32, arithmetic in
        FP64 ---- */
#define PTN_BATCH_MAX_DIM 128
#define PTN_BATCH_MAX_RANK 16

/* One record per problem and sweep. */
typedef struct {
  double f_value;      /* cost after the sweep's last factor update (before acceleration) */
  double f_accel;      /* cost at the extrapolated point, 0 when not tried */
  int    nnls_iters[3];
  int    accel_tried, accel_accepted;
  int    stopped;      /* the threshold test held after this sweep */
} ptn_batch_log;

typedef struct {
  int    block, grid;  /* launch shape of the sweep kernel */
  int    tuned;        /* 1 when the autotuner chose it */
  int    sweeps;       /* sweeps run (every problem stops on its own threshold test before) */
  double seconds;      /* device time of all sweeps (HIP events around the graph launches) */
  double tune_seconds; /* wall time of the autotuner (not part of seconds) */
  int    split_problems; /* problems whose MTTKRP sums are cut in pieces that several blocks share */
  int    max_pieces;     /* the largest piece count (at most 8) */
} ptn_batch_report;

/* The batch: dims [nprob][3], ranks [nprob]; X = the tensors back to back, each
   column-major FP64 (stored as FP32 on the device).  Replaces any earlier batch. */
ptn_status ptn_batch_set(ptn_handle h, int nprob, const int *dims, const int *ranks, const double *X);
/* Synthetic batch on the device: problem p = [[A_p]] (rank ranks[p], U[0,1) factors) plus
   Gaussian noise at snr_db (none when NaN), both drawn from seed + p with the generator
   of ptn_generate_tensor, then rounded to FP32. */
ptn_status ptn_batch_generate(ptn_handle h, int nprob, const int *dims, const int *ranks, uint64_t seed,
                              double snr_db);
/* U[0,1) initial factors, problem p from seed + p; packed per problem, problems back to back. */
ptn_status ptn_batch_init_factors(ptn_handle h, uint64_t seed, double *packed_out);
/* Tensor p of the batch as stored (FP32 values widened to FP64), column-major. */
ptn_status ptn_batch_get_tensor(ptn_handle h, int p, double *X_out);
/* ALS on every problem of the batch with the options of ptn_cpd (mttkrp_form and
   fuse_accel are not consulted; arith: PTN_ARITH_F64 runs the MTTKRPs on the FP64 matrix
   cores, PTN_ARITH_F32 on the FP32 matrix cores with FP64 totals every 12 k-steps, all
   else FP64 either way).  init_packed / out_packed: factors packed per problem,
   problems back to back; res: nprob records; log: NULL or [nprob][opt->max_iter].
   block / grid: launch shape of the sweep kernel (256 or 512 threads, grid <= 8192), or both 0 to let
   the autotuner time a few shapes on the first two sweeps (state restored afterwards).
   rep: NULL or the shape used and the times. */
ptn_status ptn_cpd_batch(ptn_handle h, const double *init_packed, const ptn_options *opt, double *out_packed,
                         ptn_result *res, ptn_batch_log *log, int block, int grid, ptn_batch_report *rep);

/* ---- the learned NNLS warm start (no counterpart in the reference, whose NesterovMNLS
        always starts from the previous factor).  A per-handle MLP, 2208 -> 256 -> 256 -> 64
        with ReLU, BF16 weights, FP32 accumulation: input = the upper triangle of the R x R
        Gram V (zero-padded to 64 x 64) scaled by its largest diagonal entry, the MTTKRP row
        and the previous factor row (scaled); output = the scaled NNLS solution row.  The
        network only proposes a start: the subproblem (its proximal anchor = the previous
        row) and the stopping test stay the reference's, and rows whose proposal fails the
        KKT test start from the previous row as in the reference ---- */
#define PTN_MLP_IN 2208
#define PTN_MLP_OUT 64
#define PTN_MLP_PARAMS 647744   /* W1 256x2208, b1 256, W2 256x256, b2 256, W3 64x256, b3 64 (row-major) */

/* Weights drawn He-uniform from seed (biases zero), Adam state cleared. */
ptn_status ptn_mlp_init(ptn_handle h, uint64_t seed);
ptn_status ptn_mlp_set_params(ptn_handle h, const float *params);
ptn_status ptn_mlp_get_params(ptn_handle h, float *params);
/* Y [n x 64] = network(X [n x 2208]) (row-major FP32 on the host, rounded to BF16). */
ptn_status ptn_mlp_forward(ptn_handle h, int n, const float *X, float *Y);
/* Mean squared error over the entries c < rank[i] of each row, and its gradient with
   respect to every parameter (the layout of ptn_mlp_get_params). */
ptn_status ptn_mlp_loss_grad(ptn_handle h, int n, const float *X, const float *T, const unsigned char *rank,
                             double *loss, float *grad);
/* The training set: pairs appended by ptn_cpd with mlp_harvest = 1, or given here. */
ptn_status ptn_mlp_data_clear(ptn_handle h);
ptn_status ptn_mlp_data_add(ptn_handle h, int n, const float *X, const float *T, const unsigned char *rank);
ptn_status ptn_mlp_data_size(ptn_handle h, int64_t *n);
/* Rows [first, first + n) of the training set: features (BF16 values as FP32), targets, ranks. */
ptn_status ptn_mlp_data_get(ptn_handle h, int64_t first, int n, float *X, float *T, unsigned char *rank);
/* Adam (lr, beta 0.9 / 0.999) on minibatches of `batch` rows (a multiple of 32) drawn
   without replacement (seeded shuffles) for `epochs` epochs; epoch_loss[e] = mean batch loss. */
ptn_status ptn_mlp_train(ptn_handle h, int epochs, int batch, double lr, uint64_t seed, double *epoch_loss);

#ifdef __cplusplus
}
#endif
#endif
