/*
 * ptn_oracle.c -- FP64 CPU restatement of Partensor's dense CPD-ALS hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); see ptn_oracle.h.
 * Reference citations are paths under the reference's include/ directory.
 */
#include "ptn_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PTNO_DBL_EPS 2.220446049250313e-16

static double wall_seconds(void)
{
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return 0.0;
#endif
}

/* ------------------------------------------------------------------------- */
/* options                                                                    */
/* ------------------------------------------------------------------------- */

/* DefaultValues, PARTENSOR_basic.hpp:184-198; inner cap NesterovMNLS.hpp:187. */
void ptno_default_options(ptno_options *o)
{
  memset(o, 0, sizeof(*o));
  o->max_iter        = 20;
  o->threshold_error = 1e-3;
  o->delta1          = 1e-2;
  o->delta2          = 1e-2;
  o->acceleration    = 1;
  o->accel_coeff     = 3;
  o->accel_fail      = 0;
  o->normalization   = 1;
  for (int i = 0; i < PTNO_MAX_ORDER; ++i) o->constraints[i] = PTNO_NONNEGATIVITY;
  o->mttkrp_form     = PTNO_MTTKRP_STANDARD;
  o->cost_form       = PTNO_COST_EXPLICIT;
  o->threads         = 0;
  o->max_inner       = 50;
}

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).                         */
/* The reference seeds Eigen's Random from time(NULL) (DataGeneration.hpp:60) */
/* so its draws cannot be reproduced; this counter-based generator replaces   */
/* them.  Counter = (idx_lo, idx_hi, stream, 0), key = (seed_lo, seed_hi).    */
/* ------------------------------------------------------------------------- */

void ptno_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    if (round > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void draw4(uint64_t seed, uint32_t stream, uint64_t idx, uint32_t w[4])
{
  uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), stream, 0u};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  ptno_philox4x32_10(ctr, key, w);
}

/* 53-bit uniform in [0,1) from two 32-bit words. */
static inline double u53(uint32_t lo, uint32_t hi)
{
  uint64_t v = (((uint64_t)hi << 32) | (uint64_t)lo) >> 11;
  return (double)v * 0x1.0p-53;
}

double ptno_uniform(uint64_t seed, uint32_t stream, uint64_t idx)
{
  uint32_t w[4];
  draw4(seed, stream, idx, w);
  return u53(w[0], w[1]);
}

/* Box-Muller, cosine branch only: one normal per counter. */
double ptno_normal(uint64_t seed, uint32_t stream, uint64_t idx)
{
  uint32_t w[4];
  draw4(seed, stream, idx, w);
  double u1 = u53(w[0], w[1]);
  double u2 = u53(w[2], w[3]);
  double rad = sqrt(-2.0 * log(1.0 - u1));
  return rad * cos(6.283185307179586 * u2);
}

double ptno_exponential(uint64_t seed, uint32_t stream, uint64_t idx)
{
  uint32_t w[4];
  draw4(seed, stream, idx, w);
  return -log(1.0 - u53(w[0], w[1]));
}

static int64_t prod_dims(int N, const int64_t *dims)
{
  int64_t p = 1;
  for (int k = 0; k < N; ++k) p *= dims[k];
  return p;
}

static void factor_offsets(int N, const int64_t *dims, int R, int64_t *off)
{
  int64_t o = 0;
  for (int k = 0; k < N; ++k) { off[k] = o; o += dims[k] * (int64_t)R; }
}

/* Factors drawn column-major, factor 0 first (makeFactors, DataGeneration.hpp:564-581,
   with the time seed replaced by `seed`). */
void ptno_gen_factors(int N, const int64_t *dims, int R, uint64_t seed, int dist, double *packed)
{
  int64_t total = 0;
  for (int k = 0; k < N; ++k) total += dims[k] * (int64_t)R;
  #pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < total; ++e)
    packed[e] = dist == 1 ? ptno_exponential(seed, 0u, (uint64_t)e) : ptno_uniform(seed, 0u, (uint64_t)e);
}

/* X = [[A_0,...,A_{N-1}]]: X(i_0..i_{N-1}) = sum_r prod_k A_k(i_k, r)
   (generateTensor, DataGeneration.hpp:120-128, evaluated element-wise). */
void ptno_cp_full(int N, const int64_t *dims, int R, const double *packed, double *X)
{
  int64_t off[PTNO_MAX_ORDER];
  factor_offsets(N, dims, R, off);
  int64_t total = prod_dims(N, dims);
  #pragma omp parallel for schedule(static)
  for (int64_t idx = 0; idx < total; ++idx) {
    int64_t ik[PTNO_MAX_ORDER];
    int64_t rem = idx;
    for (int k = 0; k < N; ++k) { ik[k] = rem % dims[k]; rem /= dims[k]; }
    double acc = 0.0;
    for (int r = 0; r < R; ++r) {
      double p = packed[off[0] + ik[0] + dims[0] * r];
      for (int k = 1; k < N; ++k) p *= packed[off[k] + ik[k] + dims[k] * r];
      acc += p;
    }
    X[idx] = acc;
  }
}

/* square_norm, TensorOperations.hpp:144-149 (sum of squares, fixed chunk order). */
double ptno_square_norm(int64_t n, const double *X)
{
  const int64_t chunk = 1 << 16;
  int64_t nchunks = (n + chunk - 1) / chunk;
  double *part = (double *)malloc((size_t)(nchunks > 0 ? nchunks : 1) * sizeof(double));
  #pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < nchunks; ++c) {
    int64_t b = c * chunk, e = b + chunk < n ? b + chunk : n;
    double s = 0.0;
    for (int64_t i = b; i < e; ++i) s = fma(X[i], X[i], s);
    part[c] = s;
  }
  double s = 0.0;
  for (int64_t c = 0; c < nchunks; ++c) s += part[c];
  free(part);
  return s;
}

double ptno_add_noise(int64_t n, double *X, uint64_t seed, double snr_db)
{
  double *E = (double *)malloc((size_t)n * sizeof(double));
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) E[i] = ptno_normal(seed, 1u, (uint64_t)i);
  double nx = sqrt(ptno_square_norm(n, X));
  double ne = sqrt(ptno_square_norm(n, E));
  double s = nx * pow(10.0, -snr_db / 20.0) / ne;
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) X[i] = fma(s, E[i], X[i]);
  free(E);
  return s;
}

/* ------------------------------------------------------------------------- */
/* layout: unfolding and partial Khatri-Rao                                   */
/* ------------------------------------------------------------------------- */

/* Matricization.hpp:101-124 (permute 65-75): X_(n)(i_n, j), j runs over the
   other modes in increasing order, lowest fastest. */
void ptno_unfold(int N, const int64_t *dims, const double *X, int mode, double *out)
{
  int64_t total = prod_dims(N, dims);
  for (int64_t idx = 0; idx < total; ++idx) {
    int64_t rem = idx, j = 0, stride = 1, in = 0;
    for (int k = 0; k < N; ++k) {
      int64_t i = rem % dims[k]; rem /= dims[k];
      if (k == mode) { in = i; continue; }
      j += i * stride; stride *= dims[k];
    }
    out[in + dims[mode] * j] = X[idx];
  }
}

/* PartialKhatriRao.hpp:59-227 / KhatriRao.hpp:60-86: KR = A_{N-1} (.) ... (.) A_0 without
   A_mode; row j enumerates the other modes, lowest fastest. */
void ptno_partial_khatri_rao(int N, const int64_t *dims, int R, const double *packed, int mode, double *out)
{
  int64_t off[PTNO_MAX_ORDER];
  factor_offsets(N, dims, R, off);
  int64_t J = prod_dims(N, dims) / dims[mode];
  for (int64_t j = 0; j < J; ++j) {
    int64_t ik[PTNO_MAX_ORDER], rem = j;
    for (int k = 0; k < N; ++k) {
      if (k == mode) continue;
      ik[k] = rem % dims[k]; rem /= dims[k];
    }
    for (int r = 0; r < R; ++r) {
      double p = 1.0;
      for (int k = 0; k < N; ++k)
        if (k != mode) p *= packed[off[k] + ik[k] + dims[k] * r];
      out[j + J * r] = p;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* MTTKRP                                                                     */
/* ------------------------------------------------------------------------- */

/* mttkrp(), MTTKRP.hpp:60-69: M = X_(n) * PartialKhatriRao(A, n).
   Row i_n of M sums over j in increasing order; parallel over output rows so
   the result does not depend on the thread count. */
void ptno_mttkrp_naive(int N, const int64_t *dims, int R, const double *X, const double *packed,
                       int mode, double *M)
{
  int64_t off[PTNO_MAX_ORDER], stride[PTNO_MAX_ORDER];
  factor_offsets(N, dims, R, off);
  int64_t s = 1;
  for (int k = 0; k < N; ++k) { stride[k] = s; s *= dims[k]; }
  int64_t In = dims[mode];
  int64_t J = prod_dims(N, dims) / In;
  #pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < In; ++i) {
    double acc[256];
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    int64_t ik[PTNO_MAX_ORDER];
    for (int k = 0; k < N; ++k) ik[k] = 0;
    for (int64_t j = 0; j < J; ++j) {
      int64_t lin = i * stride[mode];
      for (int k = 0; k < N; ++k) if (k != mode) lin += ik[k] * stride[k];
      double x = X[lin];
      for (int r = 0; r < R; ++r) {
        double kr = 1.0;
        for (int k = 0; k < N; ++k)
          if (k != mode) kr *= packed[off[k] + ik[k] + dims[k] * r];
        acc[r] = fma(x, kr, acc[r]);
      }
      for (int k = 0; k < N; ++k) {        /* odometer over the other modes */
        if (k == mode) continue;
        if (++ik[k] < dims[k]) break;
        ik[k] = 0;
      }
    }
    for (int r = 0; r < R; ++r) M[i + In * r] = acc[r];
  }
}

/* OpenMP blocked MTTKRP, MTTKRP.hpp:102-225.  For every block of I_first columns
   of X_(n): Kr = Hadamard of one row of each factor other than {mode, first_mode}
   (:200-211), PartialKR = A_first .* Kr (:214-217), result += X_(n)(:,block) *
   PartialKR (:219), with a thread-private result merged at the end (:34-37,194).
   The unfolding is index math instead of a copy (identical values). */
void ptno_mttkrp_blocked(int N, const int64_t *dims, int R, const double *X, const double *packed,
                         int mode, int threads, double *M)
{
  int64_t off[PTNO_MAX_ORDER], stride[PTNO_MAX_ORDER];
  factor_offsets(N, dims, R, off);
  int64_t s = 1;
  for (int k = 0; k < N; ++k) { stride[k] = s; s *= dims[k]; }
  const int first = (mode == 0) ? 1 : 0;
  const int64_t In = dims[mode], If = dims[first];
  const int64_t nblocks = prod_dims(N, dims) / In / If;
  int nth = threads > 0 ? threads : 1;
#ifdef _OPENMP
  if (threads <= 0) nth = omp_get_max_threads();
#endif
  memset(M, 0, (size_t)(In * R) * sizeof(double));
  #pragma omp parallel num_threads(nth)
  {
    double *acc = (double *)calloc((size_t)(In * R), sizeof(double));   /* row-major In x R */
    double *pkr = (double *)malloc((size_t)(If * R) * sizeof(double));  /* row-major If x R */
    double kr[256];
    #pragma omp for schedule(static)
    for (int64_t b = 0; b < nblocks; ++b) {
      /* decode the block index over the modes other than {mode, first}, lowest fastest */
      int64_t rem = b, base = 0;
      for (int r = 0; r < R; ++r) kr[r] = 1.0;
      for (int k = 0; k < N; ++k) {
        if (k == mode || k == first) continue;
        int64_t ik = rem % dims[k]; rem /= dims[k];
        base += ik * stride[k];
        for (int r = 0; r < R; ++r) kr[r] *= packed[off[k] + ik + dims[k] * r];
      }
      for (int64_t row = 0; row < If; ++row)
        for (int r = 0; r < R; ++r) pkr[row * R + r] = packed[off[first] + row + If * r] * kr[r];
      /* acc(In x R) += X_(n)(:, block) (In x If) * pkr (If x R) */
      for (int64_t i = 0; i < In; ++i) {
        double *arow = acc + i * R;
        const double *xb = X + base + i * stride[mode];
        for (int64_t row = 0; row < If; ++row) {
          double x = xb[row * stride[first]];
          const double *prow = pkr + row * R;
          for (int r = 0; r < R; ++r) arow[r] = fma(x, prow[r], arow[r]);
        }
      }
    }
    #pragma omp critical
    {
      for (int64_t i = 0; i < In; ++i)
        for (int r = 0; r < R; ++r) M[i + In * r] += acc[i * R + r];
    }
    free(acc);
    free(pkr);
  }
}

/* ------------------------------------------------------------------------- */
/* small dense pieces                                                         */
/* ------------------------------------------------------------------------- */

/* G = A^T A (Cpd.hpp:227): rows summed in order within chunks of 32, then the
   chunk sums added in order. */
void ptno_gram(int64_t I, int R, const double *A, double *G)
{
  for (int a = 0; a < R; ++a)
    for (int b = 0; b < R; ++b) {
      double g = 0.0;
      for (int64_t c = 0; c < I; c += 32) {
        double s = 0.0;
        for (int64_t i = c; i < I && i < c + 32; ++i) s = fma(A[i + I * a], A[i + I * b], s);
        g = c == 0 ? s : g + s;
      }
      G[a + R * b] = g;
    }
}

/* PartialCwiseProd.hpp:54-222 -> CwiseProd.hpp:48-63: the Grams other than `mode`
   in descending index order, right-nested: G_a .* (G_b .* (... .* G_z)). */
void ptno_partial_cwise(int N, int R, const double *grams_packed, int mode, double *V)
{
  int first = 1;
  for (int k = 0; k < N; ++k) {          /* ascending = innermost first */
    if (k == mode) continue;
    const double *G = grams_packed + (int64_t)k * R * R;
    for (int e = 0; e < R * R; ++e) V[e] = first ? G[e] : G[e] * V[e];
    first = 0;
  }
}

/* Round-robin (circle-method) slot of player k in round t for n (even) players. */
static inline int rr_slot(int k, int t, int n)
{
  return k == 0 ? 0 : ((k - 1 + t) % (n - 1)) + 1;
}

/* Parallel-order cyclic Jacobi on a symmetric R x R matrix (Golub & Van Loan,
   Matrix Computations 4th ed. 8.5; Brent-Luk round-robin ordering).  Each round
   applies n/2 disjoint rotations: rows first (J^T A), then columns (A J).  A
   rotation is skipped when |a_pq| <= eps * sqrt(|a_pp a_qq|).  a is overwritten
   with the (nearly) diagonal matrix; v (optional) accumulates eigenvectors. */
static int jacobi_rr(int R, double *a, double *v)
{
  const int n = R + (R & 1);
  double *cs = (double *)malloc((size_t)n * sizeof(double));
  double *sn = (double *)malloc((size_t)n * sizeof(double));
  int *pp = (int *)malloc((size_t)n * sizeof(int));
  int *qq = (int *)malloc((size_t)n * sizeof(int));
  int sweep;
  if (v) for (int i = 0; i < R; ++i) for (int j = 0; j < R; ++j) v[i + R * j] = (i == j) ? 1.0 : 0.0;
  for (sweep = 0; sweep < 60; ++sweep) {
    int rotated = 0;
    for (int t = 0; t < n - 1; ++t) {
      int np = 0;
      for (int k = 0; k < n / 2; ++k) {
        int p = rr_slot(k, t, n), q = rr_slot(n - 1 - k, t, n);
        if (p > q) { int tmp = p; p = q; q = tmp; }
        if (q >= R) continue;
        double apq = a[p + R * q], app = a[p + R * p], aqq = a[q + R * q];
        if (apq == 0.0 || fabs(apq) <= PTNO_DBL_EPS * sqrt(fabs(app * aqq))) continue;
        double theta = (aqq - app) / (2.0 * apq);
        double tt;
        if (fabs(theta) > 1e150) tt = 0.5 / theta;
        else {
          tt = 1.0 / (fabs(theta) + sqrt(fma(theta, theta, 1.0)));
          if (theta < 0.0) tt = -tt;
        }
        double c = 1.0 / sqrt(fma(tt, tt, 1.0));
        pp[np] = p; qq[np] = q; cs[np] = c; sn[np] = tt * c; ++np;
      }
      if (np == 0) continue;
      rotated = 1;
      for (int m = 0; m < np; ++m) {            /* rows p, q: J^T A */
        int p = pp[m], q = qq[m];
        double c = cs[m], s = sn[m];
        for (int j = 0; j < R; ++j) {
          double x = a[p + R * j], y = a[q + R * j];
          a[p + R * j] = fma(c, x, -(s * y));
          a[q + R * j] = fma(s, x, c * y);
        }
      }
      for (int m = 0; m < np; ++m) {            /* columns p, q: (J^T A) J */
        int p = pp[m], q = qq[m];
        double c = cs[m], s = sn[m];
        for (int i = 0; i < R; ++i) {
          double x = a[i + R * p], y = a[i + R * q];
          a[i + R * p] = fma(c, x, -(s * y));
          a[i + R * q] = fma(s, x, c * y);
        }
        if (v) for (int i = 0; i < R; ++i) {
          double x = v[i + R * p], y = v[i + R * q];
          v[i + R * p] = fma(c, x, -(s * y));
          v[i + R * q] = fma(s, x, c * y);
        }
      }
      for (int m = 0; m < np; ++m) { a[pp[m] + R * qq[m]] = 0.0; a[qq[m] + R * pp[m]] = 0.0; }
    }
    if (!rotated) break;
  }
  free(cs); free(sn); free(pp); free(qq);
  return sweep;
}

/* ---- extreme eigenvalues: Householder tridiagonalization + Sturm multisection ----
   Golub & Van Loan, Matrix Computations 4th ed., Alg. 8.3.1 (tridiagonalization
   by Householder reflections, v = x + sign(x0)||x|| e1) and 8.4.1 (Sturm sequence
   bisection, here 31 probe points per step).  Every sum runs in index order with
   fma; the device kernel performs the same operations in the same order. */

#define PTNO_PIVMIN_BASE 2.2250738585072014e-308

/* Balanced tree sum of x[0..w) over consecutive pairs, ((x0 + x1) + (x2 + x3)) + ...
   (w a power of two; x is overwritten). */
static double tree_sum(double *x, int w)
{
  for (int s = 1; s < w; s <<= 1)
    for (int i = 0; i < w; i += 2 * s) x[i] = x[i] + x[i + s];
  return x[0];
}

/* Householder tridiagonalization (Golub & Van Loan Alg. 8.3.1) of the symmetric
   matrix stored in Vraw (n x n, element (r, c) read from Vraw[r * n + c]).  The
   working matrix t is padded to RB = 16, 32 or 64 rows/columns of zeros; step k
   eliminates row k of t, forms p_l = sum_j t[j][l] v_j with four interleaved partial
   sums over all RB rows, and takes sigma and p.v as balanced tree sums over 64
   entries.  This operation order is the GPU kernels' (lane l = index l; one wave, or
   four waves that split the four partial sums and the rows of the rank-2 update), so
   the two agree bit for bit; for n > 64 RB = n and the trees widen. */
static void tridiagonalize(int n, const double *Vraw, double *d, double *e)
{
  const int RB = n <= 16 ? 16 : n <= 32 ? 32 : n <= 64 ? 64 : n;
  int W = 64;
  while (W < RB) W <<= 1;
  double *t = (double *)calloc((size_t)RB * RB, sizeof(double));
  double *v = (double *)calloc((size_t)W, sizeof(double)), *p = (double *)calloc((size_t)W, sizeof(double));
  double *xs = (double *)calloc((size_t)W, sizeof(double)), *tmp = (double *)calloc((size_t)W, sizeof(double));
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) t[r * RB + c] = Vraw[(int64_t)r * n + c];
  for (int k = 0; k + 2 < n; ++k) {
    for (int l = 0; l < W; ++l) {
      xs[l] = (l >= k + 2 && l < n) ? t[k * RB + l] : 0.0;
      tmp[l] = xs[l] * xs[l];
    }
    const double x0 = t[k * RB + k + 1];
    const double sigma = tree_sum(tmp, W);
    if (sigma == 0.0) { e[k] = x0; continue; }
    const double mu = sqrt(fma(x0, x0, sigma));
    const double v0 = x0 <= 0.0 ? x0 - mu : x0 + mu;
    e[k] = x0 <= 0.0 ? mu : -mu;
    const double beta = 2.0 / fma(v0, v0, sigma);
    for (int l = 0; l < W; ++l) v[l] = l == k + 1 ? v0 : xs[l];
    for (int l = 0; l < W; ++l) {
      double s4[4] = {0.0, 0.0, 0.0, 0.0};
      if (l < RB)
        for (int j = 0; j < RB; ++j) s4[j & 3] = fma(t[j * RB + l], v[j], s4[j & 3]);
      p[l] = (l > k && l < n) ? beta * ((s4[0] + s4[1]) + (s4[2] + s4[3])) : 0.0;
      tmp[l] = p[l] * v[l];
    }
    const double K = (0.5 * beta) * tree_sum(tmp, W);
    for (int l = k + 1; l < n; ++l) p[l] = fma(-K, v[l], p[l]);          /* w = p - K v */
    for (int l = k + 1; l < n; ++l)
      for (int j = k + 1; j < n; ++j) t[j * RB + l] = t[j * RB + l] - (v[l] * p[j] + p[l] * v[j]);
  }
  for (int i = 0; i < n; ++i) d[i] = t[i * RB + i];
  if (n >= 2) e[n - 2] = t[(n - 2) * RB + (n - 1)];
  free(t); free(v); free(p); free(xs); free(tmp);
}

/* number of eigenvalues of the tridiagonal (d, e2 = e^2) below x */
static int sturm_count(int n, const double *d, const double *e2, double pivmin, double x)
{
  int c = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  if (q < 0.0) ++c;
  for (int i = 1; i < n; ++i) {
    q = (d[i] - x) - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++c;
  }
  return c;
}

/* j-th smallest eigenvalue (0-based) by multisection on [lo, hi], count(lo) <= j < count(hi) */
static double sturm_find(int n, const double *d, const double *e2, double pivmin, double lo, double hi, int j)
{
  for (int it = 0; it < 64; ++it) {
    double w = hi - lo;
    double sc = fabs(lo) > fabs(hi) ? fabs(lo) : fabs(hi);
    if (w <= 2.0 * PTNO_DBL_EPS * sc || w <= pivmin) break;
    int m;
    for (m = 0; m < 31; ++m) {
      double x = fma(w, (double)(m + 1) * 0.03125, lo);
      if (sturm_count(n, d, e2, pivmin, x) > j) break;
    }
    double nlo = m > 0 ? fma(w, (double)m * 0.03125, lo) : lo;
    double nhi = m < 31 ? fma(w, (double)(m + 1) * 0.03125, lo) : hi;
    if (nlo == lo && nhi == hi) break;
    lo = nlo;
    hi = nhi;
  }
  return 0.5 * (lo + hi);
}

/* Extreme singular values of the symmetric V (ComputeSVD, NesterovMNLS.hpp:45-52):
   for a symmetric matrix they are max |lambda| and min |lambda|.  Returns the
   number of bisection targets evaluated (2 or 4). */
int ptno_sym_extremes(int R, const double *V, double *L, double *mu)
{
  const int n = R;
  double *d = (double *)malloc(sizeof(double) * n), *e = (double *)calloc((size_t)n, sizeof(double));
  double *e2 = (double *)calloc((size_t)n, sizeof(double));
  tridiagonalize(n, V, d, e);
  double emax = 1.0;
  for (int i = 0; i + 1 < n; ++i) { e2[i] = e[i] * e[i]; if (e2[i] > emax) emax = e2[i]; }
  double pivmin = PTNO_PIVMIN_BASE * emax;
  double glo = d[0], ghi = d[0];
  for (int i = 0; i < n; ++i) {
    double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0);
    double lo = d[i] - r, hi = d[i] + r;
    if (lo < glo) glo = lo;
    if (hi > ghi) ghi = hi;
  }
  double pad = 2.0 * PTNO_DBL_EPS * (fabs(glo) > fabs(ghi) ? fabs(glo) : fabs(ghi)) + pivmin;
  glo -= pad;
  ghi += pad;
  double lmin = sturm_find(n, d, e2, pivmin, glo, ghi, 0);
  double lmax = sturm_find(n, d, e2, pivmin, glo, ghi, n - 1);
  int k0 = sturm_count(n, d, e2, pivmin, 0.0);
  int targets = 2;
  double amin;
  if (k0 == 0) amin = fabs(lmin);
  else if (k0 == n) amin = fabs(lmax);
  else {
    double a1 = fabs(sturm_find(n, d, e2, pivmin, glo, ghi, k0 - 1));
    double a2 = fabs(sturm_find(n, d, e2, pivmin, glo, ghi, k0));
    amin = a1 < a2 ? a1 : a2;
    targets = 4;
  }
  double amax = fabs(lmin) > fabs(lmax) ? fabs(lmin) : fabs(lmax);
  *L = amax;
  *mu = amin;
  free(d); free(e); free(e2);
  return targets;
}

int ptno_sym_eig(int R, const double *V, double *evals, double *evecs)
{
  double *a = (double *)malloc((size_t)R * R * sizeof(double));
  memcpy(a, V, (size_t)R * R * sizeof(double));
  int sweeps = jacobi_rr(R, a, evecs);
  for (int i = 0; i < R; ++i) evals[i] = a[i + R * i];
  free(a);
  return sweeps;
}

/* GLambda, NesterovMNLS.hpp:75-91. */
void ptno_glambda(double mu, double *L, double *lambda, double *q)
{
  *q = mu / *L;
  if (1.0 / *q > 1e6)      *lambda = 10.0 * mu;
  else if (1.0 / *q > 1e3) *lambda = mu;
  else                     *lambda = mu / 10.0;
  *L += *lambda;
  mu += *lambda;
  *q = mu / *L;
}

/* UpdateAlpha, NesterovMNLS.hpp:103-113: positive root of a^2' ... */
double ptno_update_alpha(double alpha, double q)
{
  double a = 1.0;
  double b = alpha * alpha - q;
  double c = -alpha * alpha;
  double D = b * b - 4.0 * a * c;
  return (-b + sqrt(D)) / 2.0;
}

/* NesterovMNLS (dense), NesterovMNLS.hpp:181-234.  Z is R x R, W and A are I x R
   column-major; A is the start point and the proximal anchor (res), and receives
   the result.  The stop test is global over all I x R entries (:217).  Returns the
   number of projected-gradient steps taken. */
int ptno_nesterov_mnls(int64_t I, int R, const double *Z, const double *W, double d1, double d2,
                       int max_inner, double *A, double *L_out, double *mu_out)
{
  double L, mu, lambda, q;
  ptno_sym_extremes(R, Z, &L, &mu);
  if (L_out) *L_out = L;
  if (mu_out) *mu_out = mu;
  ptno_glambda(mu, &L, &lambda, &q);
  const double invL = 1.0 / L;

  size_t n = (size_t)(I * R);
  double *Zt = (double *)malloc((size_t)R * R * sizeof(double));
  double *Wt = (double *)malloc(n * sizeof(double));
  double *Ac = (double *)malloc(n * sizeof(double));
  double *An = (double *)malloc(n * sizeof(double));
  double *Y  = (double *)malloc(n * sizeof(double));
  double *G  = (double *)malloc(n * sizeof(double));
  memcpy(Zt, Z, (size_t)R * R * sizeof(double));
  for (int r = 0; r < R; ++r) Zt[r + R * r] = Z[r + R * r] + lambda;       /* _mat1 += lambda*I */
  for (size_t e = 0; e < n; ++e) Wt[e] = W[e] + lambda * A[e];             /* _mat2 += lambda*res */
  memcpy(Ac, A, n * sizeof(double));
  memcpy(Y, A, n * sizeof(double));
  double alpha = 1.0;
  int iter = 0;
  for (;;) {
    double gmax = 0.0, gmin = INFINITY;
    for (int64_t i = 0; i < I; ++i)
      for (int r = 0; r < R; ++r) {
        /* (Y Zt)(i, r): four interleaved partial sums over k mod 4, combined pairwise */
        double s4[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < R; ++k) s4[k & 3] = fma(Y[i + I * k], Zt[k + R * r], s4[k & 3]);
        double s = (s4[0] + s4[1]) + (s4[2] + s4[3]);
        double g = s - Wt[i + I * r];                                      /* grad = Y Zt - Wt */
        G[i + I * r] = g;
        double gy = fabs(g * Y[i + I * r]);
        if (gy > gmax) gmax = gy;
        if (g < gmin) gmin = g;
      }
    if ((gmax <= d1 && gmin >= -d2) || iter >= max_inner) break;
    for (size_t e = 0; e < n; ++e) {
      /* G / L as RN(G / L): q0 = G * RN(1/L) and one correction with the exact residual
         G - q0 L (Markstein's theorem), the GPU kernel's sequence */
      double q0 = G[e] * invL;
      double t = Y[e] - fma(fma(-q0, L, G[e]), invL, q0);
      An[e] = t > 0.0 ? t : 0.0;                                           /* cwiseMax(0) */
    }
    double na = ptno_update_alpha(alpha, q);
    double beta = alpha * (1.0 - alpha) / (alpha * alpha + na);
    for (size_t e = 0; e < n; ++e) Y[e] = (1.0 + beta) * An[e] - beta * Ac[e];
    double *tmp = Ac; Ac = An; An = tmp;
    alpha = na;
    ++iter;
  }
  memcpy(A, Ac, n * sizeof(double));
  free(Zt); free(Wt); free(Ac); free(An); free(Y); free(G);
  return iter;
}

/* Inverse by LU with partial pivoting (what Eigen's dynamic .inverse() uses). */
int ptno_inverse(int R, const double *V, double *Vinv)
{
  double *lu = (double *)malloc((size_t)R * R * sizeof(double));
  int *piv = (int *)malloc((size_t)R * sizeof(int));
  memcpy(lu, V, (size_t)R * R * sizeof(double));
  int ok = 1;
  for (int k = 0; k < R; ++k) {
    int p = k;
    for (int i = k + 1; i < R; ++i) if (fabs(lu[i + R * k]) > fabs(lu[p + R * k])) p = i;
    piv[k] = p;
    if (lu[p + R * k] == 0.0) { ok = 0; break; }
    if (p != k) for (int j = 0; j < R; ++j) { double t = lu[k + R * j]; lu[k + R * j] = lu[p + R * j]; lu[p + R * j] = t; }
    for (int i = k + 1; i < R; ++i) {
      lu[i + R * k] /= lu[k + R * k];
      for (int j = k + 1; j < R; ++j) lu[i + R * j] = fma(-lu[i + R * k], lu[k + R * j], lu[i + R * j]);
    }
  }
  if (ok) {
    for (int c = 0; c < R; ++c) {
      double *x = Vinv + R * c;
      for (int i = 0; i < R; ++i) x[i] = (i == c) ? 1.0 : 0.0;
      for (int k = 0; k < R; ++k) if (piv[k] != k) { double t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
      for (int i = 0; i < R; ++i) for (int k = 0; k < i; ++k) x[i] = fma(-lu[i + R * k], x[k], x[i]);
      for (int i = R - 1; i >= 0; --i) {
        for (int k = i + 1; k < R; ++k) x[i] = fma(-lu[i + R * k], x[k], x[i]);
        x[i] /= lu[i + R * i];
      }
    }
  }
  free(lu); free(piv);
  return ok ? 0 : -1;
}

/* ------------------------------------------------------------------------- */
/* Normalize, Normalize.hpp:92-174 (same control flow as the dimension-tree   */
/* overload :200-303).  Quirks kept: only factors 0..N-2 other than the       */
/* weight factor are rescaled, the last factor is never touched, and the zero */
/* test looks only at the last lambda_fac entry.  Grams are rescaled by       */
/* quotients, not recomputed.                                                 */
/* ------------------------------------------------------------------------- */
void ptno_normalize(int N, const int64_t *dims, int R, int wf, double *grams, double *packed)
{
  const int last = N - 1;
  int64_t off[PTNO_MAX_ORDER];
  factor_offsets(N, dims, R, off);
  double *lam = (double *)malloc((size_t)(N - 1) * R * sizeof(double));   /* lambda_fac */
  double nf[PTNO_MAX_ORDER];
  int pass = 0;
  for (int i = 0; i < N; ++i) {
    if (i != wf) {
      const double *G = grams + (int64_t)i * R * R;
      for (int r = 0; r < R; ++r) lam[(i - pass) * R + r] = G[r + R * r];
    } else pass = 1;
  }
  double cumul = 1.0;
  for (int i = 0; i < R; ++i) {
    int nonzero = 1;
    for (int j = 0; j < last; ++j) {
      nf[j] = sqrt(lam[j * R + i]);
      nonzero = !(nf[j] == 0.0);
    }
    if (!nonzero) {
      for (int j = 0; j < last; ++j) lam[j * R + i] = 1.0;
    } else {
      pass = 0;
      for (int j = 0; j < last; ++j) {
        if (j == wf) { pass = 1; continue; }
        double inv = 1.0 / nf[j - pass];
        double *col = packed + off[j] + dims[j] * i;
        for (int64_t e = 0; e < dims[j]; ++e) col[e] *= inv;
        cumul *= nf[j - pass];
        lam[(j - pass) * R + i] = nf[j - pass];
      }
      double *wcol = packed + off[wf] + dims[wf] * i;
      for (int64_t e = 0; e < dims[wf]; ++e) wcol[e] *= cumul;
      cumul = 1.0;
    }
  }
  double *nmw = (double *)malloc((size_t)R * R * sizeof(double));
  for (int e = 0; e < R * R; ++e) nmw[e] = 1.0;
  pass = 0;
  for (int i = 0; i < last; ++i) {
    if (i == wf) { pass = 1; continue; }
    const double *l = lam + (i - pass) * R;
    double *G = grams + (int64_t)i * R * R;
    for (int b = 0; b < R; ++b)
      for (int a = 0; a < R; ++a) {
        double nm = l[a] * l[b];
        nmw[a + R * b] = nmw[a + R * b] * nm;
        G[a + R * b] = G[a + R * b] / nm;
      }
  }
  double *Gw = grams + (int64_t)wf * R * R;
  for (int e = 0; e < R * R; ++e) Gw[e] = Gw[e] * nmw[e];
  free(nmw);
  free(lam);
}

/* ------------------------------------------------------------------------- */
/* costs                                                                      */
/* ------------------------------------------------------------------------- */

/* ||X_(N-1) - A_{N-1} KR^T||_F (Cpd.hpp:131-135,145-149; CpdOpenMP.hpp:121-125). */
double ptno_cost_explicit(int N, const int64_t *dims, int R, const double *X, const double *packed)
{
  int64_t off[PTNO_MAX_ORDER];
  factor_offsets(N, dims, R, off);
  const int last = N - 1;
  const int64_t Il = dims[last];
  const int64_t J = prod_dims(N, dims) / Il;            /* columns of X_(N-1) = slab size */
  double *part = (double *)malloc((size_t)Il * sizeof(double));
  #pragma omp parallel for schedule(dynamic, 1)
  for (int64_t il = 0; il < Il; ++il) {
    double s = 0.0, kr[256];
    int64_t ik[PTNO_MAX_ORDER];
    for (int k = 0; k < last; ++k) ik[k] = 0;
    for (int64_t j = 0; j < J; ++j) {
      for (int r = 0; r < R; ++r) {
        double p = packed[off[0] + ik[0] + dims[0] * r];
        for (int k = 1; k < last; ++k) p *= packed[off[k] + ik[k] + dims[k] * r];
        kr[r] = p;
      }
      double rec = 0.0;
      for (int r = 0; r < R; ++r) rec = fma(packed[off[last] + il + Il * r], kr[r], rec);
      double d = X[j + J * il] - rec;
      s = fma(d, d, s);
      for (int k = 0; k < last; ++k) { if (++ik[k] < dims[k]) break; ik[k] = 0; }
    }
    part[il] = s;
  }
  double s = 0.0;
  for (int64_t il = 0; il < Il; ++il) s += part[il];
  free(part);
  return sqrt(s);
}

static double dot_all(int64_t n, const double *a, const double *b)
{
  double s = 0.0;
  for (int64_t e = 0; e < n; ++e) s = fma(a[e], b[e], s);
  return s;
}

/* Gram form: sqrt(||X||^2 - 2 <M_last, A_last> + <V_last, G_last>)
   (CpdCUDA.hpp:137-142; CpdDimTree.hpp:156-160 with trace(M A^T) = <M, A>). */
double ptno_cost_gram(int64_t I_last, int R, double frob_sq, const double *M_last, const double *A_last,
                      const double *V_last, const double *G_last)
{
  double ma = dot_all(I_last * R, M_last, A_last);
  double vg = dot_all((int64_t)R * R, V_last, G_last);
  return sqrt(frob_sq - 2.0 * ma + vg);
}

/* ------------------------------------------------------------------------- */
/* dimension tree (N = 3, 4)                                                  */
/* Tree of ExprTree::Create (DimTrees.hpp:541-628): root splits at ceil(N/2). */
/* N=3: left {0,1} = X x_2 A_2; leaves TTV with A_1 / A_0; right {2} =        */
/*      (X x_0 A_0) TTV A_1.                                                  */
/* N=4: left {0,1} = (X x_3 A_3) TTV A_2; leaves TTV A_1 / A_0;               */
/*      right {2,3} = (X x_0 A_0) TTV A_1; leaves TTV A_3 / A_2.              */
/* Node Gramians are Hadamard products accumulated along the path, left to    */
/* right (TreeMode_N_Product :668-713, TTVs_util :732-800).                   */
/* Intermediates are stored r-slowest: Y[r][...] with the lowest mode fastest.*/
/* ------------------------------------------------------------------------- */

typedef struct {
  int N, R;
  int64_t dims[PTNO_MAX_ORDER];
  double *YL, *YR;          /* left / right node tensors (R x 2 modes)              */
  double *gL, *gR;          /* node Gramians                                         */
} dimtree_state;

/* Y(r, a, b) = sum_{c} X(a, b, c) F(c, r) for X viewed as (A x B x C) column-major. */
static void ttm_last(int64_t A, int64_t B, int64_t C, int R, const double *X, const double *F, double *Y)
{
  #pragma omp parallel for schedule(static)
  for (int64_t ab = 0; ab < A * B; ++ab) {
    for (int r = 0; r < R; ++r) {
      double s = 0.0;
      for (int64_t c = 0; c < C; ++c) s = fma(X[ab + A * B * c], F[c + C * r], s);
      Y[ab + A * B * r] = s;
    }
  }
}

/* Y(r, b) = sum_a X(a, b) F(a, r): contraction over the fastest mode. */
static void ttm_first(int64_t A, int64_t B, int R, const double *X, const double *F, double *Y)
{
  #pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < B; ++b) {
    for (int r = 0; r < R; ++r) {
      double s = 0.0;
      for (int64_t a = 0; a < A; ++a) s = fma(X[a + A * b], F[a + A * r], s);
      Y[b + B * r] = s;
    }
  }
}

/* TTV over the middle index: out(r, a, c) = sum_b Y(r, a, b, c) f_r(b), Y viewed (A x B x C) per r. */
static void ttv_mid(int64_t A, int64_t B, int64_t C, int R, const double *Y, const double *F, double *out)
{
  for (int r = 0; r < R; ++r) {
    const double *Yr = Y + A * B * C * r;
    for (int64_t c = 0; c < C; ++c)
      for (int64_t a = 0; a < A; ++a) {
        double s = 0.0;
        for (int64_t b = 0; b < B; ++b) s = fma(Yr[a + A * b + A * B * c], F[b + B * r], s);
        out[a + A * c + A * C * r] = s;
      }
  }
}

static void hadamard(int R, const double *a, const double *b, double *out)
{
  for (int e = 0; e < R * R; ++e) out[e] = a[e] * b[e];
}

/* MTTKRP of `mode` through the tree, updating the cached node it needs, as
   UpdateTree (DimTrees.hpp:862-928) does in an ALS sweep.  V receives the leaf
   Gramian; M the MTTKRP (I_mode x R col-major). */
static void dimtree_mode(dimtree_state *s, const double *X, const double *packed, const double *grams,
                         int mode, double *M, double *V)
{
  const int N = s->N, R = s->R;
  const int64_t *d = s->dims;
  int64_t off[PTNO_MAX_ORDER];
  factor_offsets(N, d, R, off);
  const double *A0 = packed + off[0], *A1 = packed + off[1], *A2 = packed + off[2];
  const double *G0 = grams, *G1 = grams + R * R, *G2 = grams + 2 * R * R;
  if (N == 3) {
    if (mode == 0) {
      ttm_last(d[0], d[1], d[2], R, X, A2, s->YL);             /* left node: X x_2 A_2 */
      memcpy(s->gL, G2, sizeof(double) * R * R);
    }
    if (mode == 0) {                                             /* leaf 0: TTV over i1 with A_1 */
      ttv_mid(d[0], d[1], 1, R, s->YL, A1, M);
      hadamard(R, s->gL, G1, V);
    } else if (mode == 1) {                                      /* leaf 1: TTV over i0 with A_0 */
      ttv_mid(1, d[0], d[1], R, s->YL, A0, M);
      hadamard(R, s->gL, G0, V);
    } else {                                                     /* right node {2} */
      ttm_first(d[0], d[1] * d[2], R, X, A0, s->YR);             /* R x I1 x I2 */
      memcpy(s->gR, G0, sizeof(double) * R * R);
      ttv_mid(1, d[1], d[2], R, s->YR, A1, M);
      hadamard(R, s->gR, G1, V);
    }
    return;
  }
  /* N == 4 */
  const double *A3 = packed + off[3];
  const double *G3 = grams + 3 * R * R;
  if (mode == 0) {
    double *Z = (double *)malloc(sizeof(double) * d[0] * d[1] * d[2] * R);
    ttm_last(d[0] * d[1], d[2], d[3], R, X, A3, Z);              /* X x_3 A_3: R x I0 x I1 x I2 */
    ttv_mid(d[0] * d[1], d[2], 1, R, Z, A2, s->YL);              /* TTV over i2 with A_2 */
    free(Z);
    hadamard(R, G3, G2, s->gL);
    ttv_mid(d[0], d[1], 1, R, s->YL, A1, M);
    hadamard(R, s->gL, G1, V);
  } else if (mode == 1) {
    ttv_mid(1, d[0], d[1], R, s->YL, A0, M);
    hadamard(R, s->gL, G0, V);
  } else if (mode == 2) {
    double *Z = (double *)malloc(sizeof(double) * d[1] * d[2] * d[3] * R);
    ttm_first(d[0], d[1] * d[2] * d[3], R, X, A0, Z);            /* R x I1 x I2 x I3 */
    ttv_mid(1, d[1], d[2] * d[3], R, Z, A1, s->YR);              /* TTV over i1: R x I2 x I3 */
    free(Z);
    hadamard(R, G0, G1, s->gR);
    ttv_mid(d[2], d[3], 1, R, s->YR, A3, M);                     /* leaf 2: TTV over i3 */
    hadamard(R, s->gR, G3, V);
  } else {
    ttv_mid(1, d[2], d[3], R, s->YR, A2, M);                     /* leaf 3: TTV over i2 */
    hadamard(R, s->gR, G2, V);
  }
}

static int dimtree_init(dimtree_state *s, int N, const int64_t *dims, int R)
{
  if (N != 3 && N != 4) return -1;
  s->N = N; s->R = R;
  for (int k = 0; k < N; ++k) s->dims[k] = dims[k];
  s->YL = (double *)malloc(sizeof(double) * dims[0] * dims[1] * R);
  s->YR = (double *)malloc(sizeof(double) * (N == 3 ? dims[1] * dims[2] : dims[2] * dims[3]) * R);
  s->gL = (double *)malloc(sizeof(double) * R * R);
  s->gR = (double *)malloc(sizeof(double) * R * R);
  return 0;
}

static void dimtree_free(dimtree_state *s)
{
  free(s->YL); free(s->YR); free(s->gL); free(s->gR);
}

int ptno_mttkrp_dimtree_all(int N, const int64_t *dims, int R, const double *X, const double *packed, double *M_packed)
{
  dimtree_state s;
  if (dimtree_init(&s, N, dims, R)) return -1;
  double *grams = (double *)calloc((size_t)N * R * R, sizeof(double));
  double *V = (double *)malloc(sizeof(double) * R * R);
  int64_t off[PTNO_MAX_ORDER];
  factor_offsets(N, dims, R, off);
  for (int m = 0; m < N; ++m) dimtree_mode(&s, X, packed, grams, m, M_packed + off[m], V);
  free(V); free(grams);
  dimtree_free(&s);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* ALS driver                                                                 */
/* ------------------------------------------------------------------------- */

static void mttkrp_form(const ptno_options *opt, int N, const int64_t *dims, int R, const double *X,
                        const double *packed, int mode, double *M)
{
  if (opt->mttkrp_form == PTNO_MTTKRP_BLOCKED)
    ptno_mttkrp_blocked(N, dims, R, X, packed, mode, opt->threads, M);
  else
    ptno_mttkrp_naive(N, dims, R, X, packed, mode, M);
}

/* The constraint switch of update_factor (Cpd.hpp:194-225) and the Gram of Cpd.hpp:227. */
static int update_one(int constraint, int64_t I, int R, const double *V, const double *M, double d1, double d2,
                      int max_inner, double *A, double *G, int *fb_out, double *L_out, double *mu_out)
{
  int it = 0;
  switch (constraint) {
    case PTNO_UNCONSTRAINED: {                         /* A = M V^{-1} */
      double *Vi = (double *)malloc(sizeof(double) * R * R);
      ptno_inverse(R, V, Vi);
      for (int64_t i = 0; i < I; ++i)
        for (int c = 0; c < R; ++c) {
          double s = 0.0;
          for (int k = 0; k < R; ++k) s = fma(M[i + I * k], Vi[k + R * c], s);
          A[i + I * c] = s;
        }
      free(Vi);
      break;
    }
    case PTNO_NONNEGATIVITY: {
      double *old = (double *)malloc(sizeof(double) * I * R);
      memcpy(old, A, sizeof(double) * I * R);
      it = ptno_nesterov_mnls(I, R, V, M, d1, d2, max_inner, A, L_out, mu_out);
      /* zero-column fallback, Cpd.hpp:206-207 */
      double mincol = INFINITY;
      for (int c = 0; c < R; ++c) {
        double s = 0.0;
        for (int64_t i = 0; i < I; ++i) s += fabs(A[i + I * c]);
        if (s < mincol) mincol = s;
      }
      int fb = 0;
      if (mincol == 0.0) {
        fb = 1;
        for (int64_t e = 0; e < I * R; ++e) A[e] = 0.9 * A[e] + 0.1 * old[e];
      }
      *fb_out = fb;
      free(old);
      break;
    }
    case PTNO_ORTHOGONALITY: {                         /* A = M (M^T M)^{-1/2} */
      double *MtM = (double *)malloc(sizeof(double) * R * R);
      double *ev = (double *)malloc(sizeof(double) * R);
      double *Q = (double *)malloc(sizeof(double) * R * R);
      double *T = (double *)malloc(sizeof(double) * R * R);
      ptno_gram(I, R, M, MtM);
      ptno_sym_eig(R, MtM, ev, Q);
      for (int a = 0; a < R; ++a)
        for (int b = 0; b < R; ++b) {
          double s = 0.0;
          for (int k = 0; k < R; ++k) s = fma(Q[a + R * k] * sqrt(1.0 / ev[k]), Q[b + R * k], s);
          T[a + R * b] = s;
        }
      for (int64_t i = 0; i < I; ++i)
        for (int c = 0; c < R; ++c) {
          double s = 0.0;
          for (int k = 0; k < R; ++k) s = fma(M[i + I * k], T[k + R * c], s);
          A[i + I * c] = s;
        }
      free(MtM); free(ev); free(Q); free(T);
      break;
    }
    default: break;                                     /* sparsity / constant: unchanged */
  }
  ptno_gram(I, R, A, G);
  return it;
}

/* update_factor, Cpd.hpp:189-228 (dimension tree: CpdDimTree.hpp:199-239). */
static void update_factor(const ptno_options *opt, int idx, int64_t I, int R, const double *V,
                          const double *M, double *A, double *G, ptno_sweep_log *lg)
{
  int fb = 0;
  double L = 0, mu = 0;
  int it = update_one(opt->constraints[idx], I, R, V, M, opt->delta1, opt->delta2, opt->max_inner, A, G, &fb,
                      &L, &mu);
  if (lg && opt->constraints[idx] == PTNO_NONNEGATIVITY) {
    lg->nnls_iters[idx] = it; lg->zero_col_fallback[idx] = fb; lg->lipschitz[idx] = L; lg->mu[idx] = mu;
  }
}

int ptno_update_factor(int constraint, int64_t I, int R, const double *V, const double *M, double d1,
                       double d2, int max_inner, double *A, double *G, int *fallback)
{
  double L = 0, mu = 0;
  int fb = 0;
  int it = update_one(constraint, I, R, V, M, d1, d2, max_inner, A, G, &fb, &L, &mu);
  if (fallback) *fallback = fb;
  return it;
}

int ptno_cpd(int N, const int64_t *dims, int R, const double *X, const double *init_packed,
             const ptno_options *opt, double *out_packed, ptno_status *st,
             ptno_sweep_log *log, int log_cap)
{
  if (N < 3 || N > PTNO_MAX_ORDER || R < 1 || R > 256) return -1;
  if (opt->mttkrp_form == PTNO_MTTKRP_DIMTREE && N != 3 && N != 4) return -2;
  const int last = N - 1;
  int64_t off[PTNO_MAX_ORDER];
  factor_offsets(N, dims, R, off);
  int64_t fsz = off[last] + dims[last] * R;
  int64_t total = prod_dims(N, dims);

  double *A     = out_packed;
  double *old   = (double *)malloc(sizeof(double) * fsz);
  double *nrm   = (double *)malloc(sizeof(double) * fsz);
  double *acc   = (double *)malloc(sizeof(double) * fsz);
  double *grams = (double *)malloc(sizeof(double) * N * R * R);
  double *agr   = (double *)malloc(sizeof(double) * N * R * R);
  double *V     = (double *)malloc(sizeof(double) * R * R);
  int64_t maxI = 0;
  for (int k = 0; k < N; ++k) if (dims[k] > maxI) maxI = dims[k];
  double *M     = (double *)malloc(sizeof(double) * maxI * R);
  memcpy(A, init_packed, sizeof(double) * fsz);

  dimtree_state tree;
  int use_tree = opt->mttkrp_form == PTNO_MTTKRP_DIMTREE;
  if (use_tree) dimtree_init(&tree, N, dims, R);

  for (int k = 0; k < N; ++k) ptno_gram(dims[k], R, A + off[k], grams + (int64_t)k * R * R);

  /* choose_normilization_factor, Normalize.hpp:53-67 */
  int all_orth = 1, wf = 0;
  if (opt->normalization)
    for (int k = 0; k < N; ++k)
      if (opt->constraints[k] != PTNO_ORTHOGONALITY) { all_orth = 0; wf = k; break; }

  st->frob_sq = ptno_square_norm(total, X);
  st->f_value = ptno_cost_explicit(N, dims, R, X, A);               /* cost_function_init */
  st->rel_cost = st->f_value / sqrt(st->frob_sq);
  st->ao_iter = 0;
  st->accel_coeff = opt->accel_coeff;
  st->accel_fail = opt->accel_fail;

  for (;;) {
    st->ao_iter++;
    ptno_sweep_log lrec;
    memset(&lrec, 0, sizeof(lrec));
    lrec.t_start = wall_seconds();
    for (int i = 0; i < N; ++i) {
      if (use_tree) dimtree_mode(&tree, X, A, grams, i, M, V);
      else {
        mttkrp_form(opt, N, dims, R, X, A, i, M);
        ptno_partial_cwise(N, R, grams, i, V);
      }
      update_factor(opt, i, dims[i], R, V, M, A + off[i], grams + (int64_t)i * R * R, &lrec);
      if (i == last) {
        if (opt->cost_form == PTNO_COST_GRAM)
          st->f_value = ptno_cost_gram(dims[last], R, st->frob_sq, M, A + off[last], V, grams + (int64_t)last * R * R);
        else
          st->f_value = ptno_cost_explicit(N, dims, R, X, A);
      }
    }
    st->rel_cost = st->f_value / sqrt(st->frob_sq);
    lrec.f_value = st->f_value;
    lrec.rel_cost = st->rel_cost;
    if (opt->normalization && !all_orth) ptno_normalize(N, dims, R, wf, grams, A);

    int stop = st->rel_cost < opt->threshold_error || st->ao_iter >= opt->max_iter;
    if (!stop && opt->acceleration) {
      memcpy(nrm, A, sizeof(double) * fsz);
      if (st->ao_iter > 1) {
        /* line_search_accel, Cpd.hpp:247-278 (cost :162-169) */
        double step = pow((double)(st->ao_iter + 1), 1.0 / (double)st->accel_coeff);
        for (int64_t e = 0; e < fsz; ++e) acc[e] = old[e] + step * (A[e] - old[e]);
        for (int k = 0; k < N; ++k) ptno_gram(dims[k], R, acc + off[k], agr + (int64_t)k * R * R);
        if (opt->mttkrp_form == PTNO_MTTKRP_BLOCKED)
          ptno_mttkrp_blocked(N, dims, R, X, acc, last, opt->threads, M);
        else
          ptno_mttkrp_naive(N, dims, R, X, acc, last, M);
        ptno_partial_cwise(N, R, agr, last, V);
        double vg = dot_all((int64_t)R * R, V, agr + (int64_t)last * R * R);
        double ma = dot_all(dims[last] * R, M, acc + off[last]);
        double fa = sqrt(st->frob_sq + vg - 2.0 * ma);
        lrec.accel_tried = 1;
        lrec.f_accel = fa;
        int accept = st->f_value > fa;
        if (opt->force_accept && st->ao_iter <= opt->force_n && opt->force_accept[st->ao_iter - 1] >= 0) {
          int forced = opt->force_accept[st->ao_iter - 1] != 0;
          lrec.accel_overridden = forced != accept;
          accept = forced;
        }
        if (accept) {
          memcpy(A, acc, sizeof(double) * fsz);
          memcpy(grams, agr, sizeof(double) * N * R * R);
          st->f_value = fa;
          lrec.accel_accepted = 1;
        } else {
          st->accel_fail++;
        }
        if (st->accel_fail == 5) { st->accel_fail = 0; st->accel_coeff++; }
      }
      memcpy(old, nrm, sizeof(double) * fsz);
    }
    lrec.t_end = wall_seconds();
    if (log && st->ao_iter - 1 < log_cap) log[st->ao_iter - 1] = lrec;
    if (stop) break;
  }

  if (use_tree) dimtree_free(&tree);
  free(old); free(nrm); free(acc); free(grams); free(agr); free(V); free(M);
  return 0;
}
