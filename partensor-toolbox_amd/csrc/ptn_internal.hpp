// Internal declarations shared by the HIP kernels and the C-ABI layer.
//
// Device layouts (all chosen for the GPU, converted at the ABI boundary):
//   tensor X      column-major, mode 0 fastest (as the reference's Eigen Tensor)
//   factor A_k    row-major I_k x ldr doubles (ldr = R rounded up to 16, padding 0),
//                 so one factor row is contiguous for the KR products and NNLS rows
//   MTTKRP M_n    row-major I_n x ldr doubles
//   Gram / V      row-major R x R doubles (symmetric)
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ptn {

constexpr int kMaxOrder = 8;
constexpr int kMaxRank = 64;        // one NNLS row per 64-lane wavefront
constexpr int kMaxInner = 255;      // satisfaction masks hold 4 x 64 bits

// Mixed-radix description of an index set enumerated lowest-mode-fastest,
// with the factor row products attached to it (a Khatri-Rao row generator).
struct KrSpec {
  int nm;                          // number of modes in the set
  int64_t dim[kMaxOrder];          // their sizes
  int64_t xstride[kMaxOrder];      // their strides in the tensor X (elements)
  const double *fac[kMaxOrder];    // row-major factor of each mode (nullptr: not multiplied)
  const float *fac32[kMaxOrder];   // FP32 copies (bf16x3 / FP32 kernels)
};

// Per-mode NNLS constants produced on the device by the eigen kernel.
struct NnlsParams {
  double L0, mu0;       // extreme singular values of V (before GLambda)
  double L, lambda, q;  // after GLambda
  int iters;            // projected-gradient steps taken (k*)
  int fallback;         // zero-column blend applied
  int warm_rows;        // rows started from the warm-start network's proposal
  int warm_steps;       // Nesterov steps run from the network's proposals before the KKT test
};

// ---- batched small CPDs (batch_kernels.hip) ----
// One order-3 problem of a batch: sizes, rank (<= 16), where its tensor, factor rows
// (rows of 16 doubles, factors 0, 1, 2 back to back) and Grams (3 x 16 x 16) live.
// A problem's MTTKRPs are summed in V "virtual members" (V = 1: one block's waves, as
// for small problems; V > 1: the (band, k-step) space cut in V contiguous pieces, each
// summed by a block's waves, the pieces added in order), so its result does not depend
// on how many blocks run it: G blocks (members) each take V / G pieces, exchange them
// through global memory and then all run the rest of the sweep on identical data; member
// 0 keeps the problem's state, members 1 .. G-1 private copies of it (h_off, hg_off).
struct BatchDesc {
  int I[3];
  int R;
  int rowoff[3];
  int V;                       // virtual members of the MTTKRP sums
  int64_t x_off, f_off, g_off;
  double frob;                 // ||X||^2, set by batch_init
  int G;                       // blocks running the problem (1 .. V)
  int pad;
  int64_t h_off, hg_off;       // the helpers' factor rows / Grams (helper g at (g - 1) x size)
  int64_t mx_off;              // exchange area: [2 parities][V][128 x 16] doubles
};
// ALS state of one problem, kept on the device between sweeps.
struct BatchProb {
  double f, f_accel;
  int accel_coeff, accel_fail;
  int done;                    // stopped on the threshold
  int sweeps;                  // sweeps completed
  int tried, accepted;         // line search of the last sweep
  int nnls_iters[3];
  int nx;                      // group exchanges done (members wait for G x nx arrivals)
};
// One record per problem and sweep (layout of ptn_batch_log in the C ABI).
struct BatchLog {
  double f_value, f_accel;
  int nnls_iters[3];
  int accel_tried, accel_accepted, stopped;
};
struct BatchArgs {
  static constexpr int kMaxThreads = 512;
  int nprob;
  BatchDesc *desc;
  BatchProb *st;
  const float *X;
  double *fac, *old, *acc, *grams;   // factor rows (current, previous sweep, extrapolated), Grams
  const int *order;            // problems, largest first
  int *queue;                  // one work counter per sweep (zero-initialised)
  const int *sweep_ctr;        // the sweep being run (1-based), bumped after each sweep
  int max_iter;
  double thr, d1, d2;
  int max_inner, acceleration, normalization;
  const double *pow_table;     // (sweep + 1)^(1 / coeff) at [sweep * pow_stride + coeff]
  int pow_stride;
  BatchLog *log;               // [nprob][max_iter] (optional)
  int arith_f32;               // 1: MTTKRP on the FP32 matrix cores (FP64 totals), 0: FP64 matrix cores
  int64_t xcopy;               // X + n xcopy: the tensors with mode n fastest ((i1, i0, i2), (i2, i0, i1) for n = 1, 2)
  int nitems;                  // queue entries: one per (problem, member), order[q] = 8 p + member
  double *hfac, *hold, *hacc, *hgrams;   // helpers' private state (BatchDesc::h_off, hg_off)
  double *mx;                  // exchange areas (BatchDesc::mx_off)
  int *gctr;                   // per problem: arrivals at the group's exchanges
  int *err;                    // set when a group wait runs out (the run is then reported failed)
  double *pscratch;            // 256-thread blocks: partial MTTKRP bands, kBatchPartDoubles256 per block
};
constexpr int kBatchMaxMembers = 8;
constexpr int64_t kBatchPartDoubles256 = 4 * 2 * 32 * 16;
// copies every problem's state to its helpers' private copies (after set-up or a restore)
void batch_sync_helpers(hipStream_t s, const BatchArgs &a);
size_t batch_lds_bytes();
void batch_init(hipStream_t s, const BatchArgs &a);
void batch_sweep(hipStream_t s, const BatchArgs &a, int block, int grid);
void batch_next_sweep(hipStream_t s, int *ctr);
// the mode-1 / mode-2 copies of one problem's tensor (BatchArgs::xcopy)
void batch_permute(hipStream_t s, const float *X, int I0, int I1, int I2, float *X1, float *X2);

// ---- update_factor's unconstrained and orthogonality branches (constraint_kernels.hip) ----
void lu_inverse(hipStream_t s, int R, const double *V, double *Vi);            // NaN when singular
void jacobi_eig(hipStream_t s, int R, const double *A, double *evals, double *Q);
void inv_sqrt_matrix(hipStream_t s, int R, const double *Q, const double *ev, double *T);
void apply_right(hipStream_t s, int64_t I, int R, int ldr, const double *M, const double *T, double *A);
void params_no_nnls(hipStream_t s, NnlsParams *p);

// ---- the NNLS warm-start network (mlp_kernels.hip) ----
// 2208 -> 256 -> 256 -> 64 with ReLU: inputs vech(V) (64 x 65 / 2) + MTTKRP row + previous row.
constexpr int kMlpR = 64, kMlpVech = kMlpR * (kMlpR + 1) / 2, kMlpIn = kMlpVech + 2 * kMlpR, kMlpH = 256;
// FP32 master vector: W1 [H x In], b1 [H], W2 [H x H], b2 [H], W3 [R x H], b3 [R]
constexpr int64_t kMlpW1 = 0, kMlpB1 = kMlpW1 + (int64_t)kMlpH * kMlpIn, kMlpW2 = kMlpB1 + kMlpH,
                  kMlpB2 = kMlpW2 + (int64_t)kMlpH * kMlpH, kMlpW3 = kMlpB2 + kMlpH,
                  kMlpB3 = kMlpW3 + (int64_t)kMlpR * kMlpH, kMlpParams = kMlpB3 + kMlpR;
// bf16 copies of W1, W2, W3 back to back
constexpr int64_t kMlpW1b = 0, kMlpW2b = (int64_t)kMlpH * kMlpIn, kMlpW3b = kMlpW2b + (int64_t)kMlpH * kMlpH,
                  kMlpWb = kMlpW3b + (int64_t)kMlpR * kMlpH;
__host__ __device__ inline int64_t mlp_bf16_index(int64_t e) {
  if (e < kMlpB1) return e;
  if (e >= kMlpW2 && e < kMlpB2) return e - kMlpW2 + kMlpW2b;
  if (e >= kMlpW3 && e < kMlpB3) return e - kMlpW3 + kMlpW3b;
  return -1;
}
void gemm_nt(hipStream_t s, int M, int N, int K, const void *A, int lda, const void *B, int ldb, const float *bias,
             const void *gate, int ldg, int relu, float *C, int ldc, void *Cb, int ldcb);
void transpose_bf16(hipStream_t s, int rows, int cols, const void *in, int ldi, void *out, int ldo);
void mlp_features(hipStream_t s, int64_t I, int R, int ldr, const double *V, const double *M, const double *A,
                  const double *Anew, void *X, float *T, unsigned char *rank, float *scale);
// C: the proposals after warm_refine Nesterov steps (null: the network's outputs Y, scale)
void warm_select(hipStream_t s, int64_t I, int R, int ldr, const double *V, NnlsParams *p, const double *M,
                 const double *A, const float *Y, const float *scale, const double *C, double kkt_tol, double *S,
                 int *count);
void mse_grad(hipStream_t s, int B, const float *Y, const float *T, const unsigned char *rank, float inv_count,
              void *dYb, float *dY, double *part, int nparts);
void colsum_f32(hipStream_t s, int M, int N, const float *G, int ldg, float *out);
void colsum_bf16(hipStream_t s, int M, int N, const void *G, int ldg, float *out);
void adam_step(hipStream_t s, int64_t n, float *w, const float *g, float *m1, float *m2, float lr, float b1, float b2,
               float eps, int step, void *wb);
void weights_bf16(hipStream_t s, int64_t n, const float *w, void *wb);
void gather_rows(hipStream_t s, int B, const int *idx, const void *X, const float *T, const unsigned char *rank,
                 void *Xb, float *Tb, unsigned char *rb);

// ---- generation (gen_kernels.hip) ----
void gen_factors(hipStream_t s, int N, const int64_t *dims, int R, uint64_t seed, int dist, double *packed);
void cp_full(hipStream_t s, int N, const int64_t *dims, int R, const double *packed_dev,
             double *x64, float *x32);
void gen_noise_norm(hipStream_t s, int64_t n, uint64_t seed, double *partials, int nparts);
void add_noise(hipStream_t s, int64_t n, uint64_t seed, const double *scale_dev, double *x64, float *x32);
void square_norm_partials(hipStream_t s, int64_t n, const double *x64, const float *x32, double *partials, int nparts);
void sum_partials(hipStream_t s, const double *partials, int nparts, double *out);
void noise_scale(hipStream_t s, const double *nx2, const double *ne2, double snr_db, double *scale);
void cast_f64_to_f32(hipStream_t s, int64_t n, const double *src, float *dst);
void cost_explicit_partials(hipStream_t s, int N, const int64_t *dims, int R, int ldr, const double *const *fac,
                            const double *x64, const float *x32, double *partials, int nparts);

// ---- layout conversion ----
// zeroes `words` 8-byte words at p with a kernel (no memset node)
void fill_zero(hipStream_t s, void *p, int64_t words);
void colmajor_to_rows(hipStream_t s, int64_t I, int R, int ldr, const double *src, double *dst);
void rows_to_colmajor(hipStream_t s, int64_t I, int R, int ldr, const double *src, double *dst);
void rows_to_f32(hipStream_t s, int64_t n, const double *src, float *dst);

// ---- MTTKRP building blocks (mttkrp_kernels.hip) ----
// out[split][m][r] = sum_{k in split} X[m + Mrows*k] * W(k)[r];  W = Khatri-Rao row of `w`.
void gemm_cm_f64(hipStream_t s, const double *x64, const float *x32, int64_t Mrows, int64_t Ktot,
                 const KrSpec &w, int ldr, int splits, double *out);
// T[c][r] = sum_{a < L} X[a + L*c] * W(a)[r]   for c < C;  W = Khatri-Rao row of `w` (a < 2^32)
void gemm_ck_f64(hipStream_t s, const double *x64, const float *x32, int64_t L, int64_t C,
                 const KrSpec &w, int ldr, double *T);
// out[i][r] (+)= sum over rows t of T whose output digit is i of T[t][r] * prod of the
// other digits' factor rows.  `t` enumerates spec modes lowest fastest; `outmode` is the
// position in spec of the output index.  Rows of T are ldt apart; T may be FP64 or FP32.
// sel (optional, device): the node starts at T + (*sel) * sel_off
void contract_rows(hipStream_t s, const double *T64, const float *T32, int64_t ldt, const KrSpec &spec,
                   int outmode, int ldr, int splits, double *partials, const int *sel = nullptr,
                   int64_t sel_off = 0);
void reduce_splits(hipStream_t s, const double *partials, int splits, int64_t n, double *out);
// Two-digit node Y (rows i0 + D0 i1, ld values per row) contracted against the other
// digit's factor F: part[split][i][ldr] for the kept digit i (outmode 0: i0, 1: i1).
// sel (optional, device): the node starts at Y + (*sel) * sel_off.
void contract2(hipStream_t s, const double *Y64, const float *Y32, int64_t ld, int64_t D0, int64_t D1, int outmode,
               const double *F, int ldr, int R, int splits, double *part, const int *sel = nullptr,
               int64_t sel_off = 0);
// contract2 of two FP32 nodes of one shape in one launch: part[split][0 / 1][i][ldr] for
// (Ya, Fa) / (Yb, Fb); reduce_splits over 2 x D_out x ldr then gives both results.
void contract2_pair(hipStream_t s, const float *Ya, const float *Yb, int64_t ld, int64_t D0, int64_t D1, int outmode,
                    const double *Fa, const double *Fb, int ldr, int R, int splits, double *part);

// FP32-stream MFMA variants (bf16x3 split products or FP32 MFMA), FP32 outputs with
// RPF = R rounded up to 32 columns.  The Khatri-Rao operand is packed once per
// contraction by pack_w (bf16 hi/lo fragments for bf16x3, FP32 rows for FP32 MFMA).
enum GemmMode { kGemmFp32 = 0, kGemmBf16x3 = 1 };
int64_t wpack_elems(int64_t K, int RPF);   // number of 16-byte fragments (per hi / lo array)
// kperm 0: fragments for gemm_cm_f32 (16-wide k steps); kperm 1: for gemm_ck_f32 (32-wide)
// w2 (optional): a second operand packed into the upper half of the RPF columns.
void pack_w(hipStream_t s, const KrSpec &w, const KrSpec *w2, int64_t K, int RPF, int ldr, int kperm, void *whi,
            void *wlo, float *wf);
void gemm_cm_f32(hipStream_t s, GemmMode mode, const float *x32, int64_t M, int64_t K, int RPF, int splits,
                 const void *whi, const void *wlo, const float *wf, float *out);
// fuse_I1 > 0: C = I1 * I2 fibers, epilogue TTV with A1 (FP64 rows, ldr) into part[I2][ceil(I1/64)][ldr]
// (reduce_chunks then gives the order-3 mode-2 MTTKRP); fuse_I1 == 0: writes T[c][RPF].
void gemm_ck_f32(hipStream_t s, GemmMode mode, const float *x32, int64_t L, int64_t C, int RPF, const void *whi,
                 const void *wlo, const float *wf, float *T, int64_t fuse_I1, const double *A1, int ldr,
                 double *part);
void reduce_chunks(hipStream_t s, const double *part, int64_t I2, int64_t wpf, int ldr, double *M);
void reduce_splits_f32(hipStream_t s, const float *partials, int splits, int64_t n, float *out);

// ---- small dense pieces (small_kernels.hip) ----
void hadamard(hipStream_t s, int R, int ng, const double *const *grams, double *V);
// partials: ceil(I/32) x R x R doubles
void gram(hipStream_t s, int64_t I, int R, int ldr, const double *A, double *G, double *partials);
void eig_extremes(hipStream_t s, int R, const double *V, NnlsParams *p);
// V = Hadamard of the ng Grams (list order), its extreme eigenvalues with GLambda, and
// the max_inner momentum coefficients: everything an NNLS update needs besides M.
void nnls_setup(hipStream_t s, int R, int ng, const double *const *grams, double *V, NnlsParams *p, int max_inner,
                double *betas);
// betas: max_inner doubles
void nnls_prep(hipStream_t s, const NnlsParams *p, int max_inner, double *betas);
// satisfaction masks of nnls_iterate: 4 words per block of 4 rows
inline int64_t nnls_mask_words(int64_t I) { return (I + 3) / 4 * 4; }
// hist: (max_inner + 1) x I x ldr doubles; mask: nnls_mask_words(I) words
// A: the previous factor rows (the proximal anchor, and the start when S is null);
// S: another start (the warm start)
void nnls_iterate(hipStream_t s, int64_t I, int R, int ldr, const double *V, const NnlsParams *p,
                  const double *betas, const double *M, const double *A, const double *S, double d1, double d2,
                  int max_inner, uint64_t *mask, double *hist, double *colsum);   // zeroes colsum for nnls_select
void nnls_select(hipStream_t s, int64_t I, int R, int ldr, const double *hist, const uint64_t *mask, int max_inner,
                 double *Anew, double *colsum, NnlsParams *p);
// nnls_select over 32-row blocks that also leave partial Grams (partials: ceil(I/32) x R x R),
// then the zero-column test of update_factor (Cpd.hpp:206-207; A = 0.9 A + 0.1 Aold when a
// column abs-sum is zero) and G = A^T A in gram's association: the end of a
// non-negative factor update in two launches.
void nnls_finish(hipStream_t s, int64_t I, int R, int ldr, const double *hist, const uint64_t *mask, int max_inner,
                 double *A, double *colsum, const double *Aold, double *G, double *partials, NnlsParams *p);
void cost_gram(hipStream_t s, int64_t nma, const double *M, const double *A, int R, const double *V, const double *G,
               const double *frob, double *f, int accel);

// up to four ranges of 4-byte words copied to mapped host memory by one kernel; with
// seq_dst the kernel then publishes seq there (after a system-scope fence), telling the
// host that the readback of the sweep is complete
struct ExportList {
  const unsigned *src[4];
  unsigned *dst[4];
  int words[4];
  int count;
  unsigned *seq_dst;
  unsigned seq;
};
void export_words(hipStream_t s, const ExportList &x);
// the line-search step fused into Normalize's row pass (k_axpy_accel's arithmetic); cur is
// the packed factor base the rows belong to, old / out / snap are indexed like it
struct AxpyArgs {
  double step_c, step_c1;
  int c;
  const int *coeff;
  const double *cur;
  double *old, *out, *snap;
};
// ex (optional): the sweep's readback, copied by the first kernel; ax (optional): the
// line-search step in the row pass
void normalize(hipStream_t s, int N, const int64_t *dims, int R, int ldr, int wf, double *const *grams,
               double *const *fac, double *scratch, const ExportList *ex = nullptr, const AxpyArgs *ax = nullptr);
// Conditional copies of an accepted line-search step: four double ranges and one
// NnlsParams (any may be empty / null).  pick: one range copied from pick_src[0] when the
// step is accepted, else from pick_src[1].  fa_out / flags_out (mapped host memory,
// optional): the accelerated cost and the four flags after the decision.
struct AccelCopy {
  const double *src[4];
  double *dst[4];
  int64_t n[4];
  const NnlsParams *prm_src;
  NnlsParams *prm_dst;
  const double *pick_src[2];
  double *pick_dst;
  int64_t pick_n;
  double *fa_out;
  int *flags_out;
  double *f_out;                             // optional: the sweep's cost
  unsigned *seq_dst;                         // optional: published last, as in ExportList
  unsigned seq;
};
void accel_decide(hipStream_t s, const double *scal, double thr, int *flags, const AccelCopy &cp);
void axpy_accel(hipStream_t s, int64_t n, double step_c, double step_c1, int c, const int *coeff, double *old,
                const double *cur, double *out, double *snap = nullptr);


}  // namespace ptn
