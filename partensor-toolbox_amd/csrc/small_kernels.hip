// R x R and row-wise pieces of an ALS update, all FP64 on the device:
// Gram Hadamard products (PartialCwiseProd.hpp:54-222), Grams (Cpd.hpp:227),
// extreme singular values of V + GLambda (NesterovMNLS.hpp:45-91), the batched
// Nesterov NNLS (NesterovMNLS.hpp:181-234), the zero-column blend
// (Cpd.hpp:206-207), Normalize (Normalize.hpp:92-174), the Gram-form costs
// (CpdCUDA.hpp:137-142, Cpd.hpp:162-169) and the acceleration extrapolation
// (Cpd.hpp:247-258).  Built with -ffp-contract=off: fused multiply-adds appear
// only where written as fma(), so the arithmetic sequence is fixed.
#include "ptn_internal.hpp"

namespace ptn {

namespace {

struct GramPtrs {
  const double *p[kMaxOrder];
};
struct MutPtrs {
  double *p[kMaxOrder];
};

constexpr double kEps = 2.220446049250313e-16;

__device__ __forceinline__ double bcast_lane(double v, int k) {
  int2 p = __builtin_bit_cast(int2, v);
  p.x = __builtin_amdgcn_readlane(p.x, k);
  p.y = __builtin_amdgcn_readlane(p.y, k);
  return __builtin_bit_cast(double, p);
}

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// Accumulate in list order: V = G_l0; V = G_lj * V (right-nested when the list is
// ascending, left-nested along a tree path when given in path order).
__global__ void k_hadamard(int R, int ng, GramPtrs G, double *V) {
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
    double v = G.p[0][e];
    for (int j = 1; j < ng; ++j) v = G.p[j][e] * v;
    V[e] = v;
  }
}

__global__ void k_gram_partials(int64_t I, int R, int ldr, const double *__restrict__ A, int64_t rchunk,
                                double *__restrict__ partials) {
  const int64_t i0 = (int64_t)blockIdx.x * rchunk;
  const int64_t i1 = i0 + rchunk < I ? i0 + rchunk : I;
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
    int a = e / R, b = e % R;
    double s = 0.0;
    for (int64_t i = i0; i < i1; ++i) s = fma(A[i * ldr + a], A[i * ldr + b], s);
    partials[(int64_t)blockIdx.x * R * R + e] = s;
  }
}

__global__ void k_gram_reduce(int R, const double *__restrict__ partials, int nparts, double *__restrict__ G) {
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += partials[(int64_t)p * R * R + e];
    G[e] = s;
  }
}

__device__ __forceinline__ int rr_slot(int k, int t, int n) { return k == 0 ? 0 : ((k - 1 + t) % (n - 1)) + 1; }

// Parallel-order cyclic Jacobi (Golub & Van Loan 8.5, Brent-Luk round robin): each
// round applies the n/2 disjoint rotations, rows first (J^T A) then columns (A J);
// a rotation is skipped when |a_pq| <= eps sqrt(|a_pp a_qq|); stops after a sweep
// with no rotation.  Singular values of the symmetric V are |eigenvalues|.
__global__ __launch_bounds__(256) void k_eig_extremes(int R, const double *__restrict__ V, NnlsParams *prm) {
  __shared__ double a[kMaxRank * kMaxRank];
  __shared__ double cs[kMaxRank / 2], sn[kMaxRank / 2];
  __shared__ int pp[kMaxRank / 2], qq[kMaxRank / 2], act[kMaxRank / 2];
  __shared__ int rotated;
  const int n = R + (R & 1);
  const int half = n / 2;
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) a[e] = V[e];
  __syncthreads();
  for (int sweep = 0; sweep < 60; ++sweep) {
    if (threadIdx.x == 0) rotated = 0;
    __syncthreads();
    for (int t = 0; t < n - 1; ++t) {
      if (threadIdx.x < half) {
        int k = threadIdx.x;
        int p = rr_slot(k, t, n), q = rr_slot(n - 1 - k, t, n);
        if (p > q) { int tmp = p; p = q; q = tmp; }
        int on = 0;
        if (q < R) {
          double apq = a[p * R + q], app = a[p * R + p], aqq = a[q * R + q];
          if (!(apq == 0.0 || fabs(apq) <= kEps * sqrt(fabs(app * aqq)))) {
            double theta = (aqq - app) / (2.0 * apq);
            double tt;
            if (fabs(theta) > 1e150) tt = 0.5 / theta;
            else {
              tt = 1.0 / (fabs(theta) + sqrt(fma(theta, theta, 1.0)));
              if (theta < 0.0) tt = -tt;
            }
            double c = 1.0 / sqrt(fma(tt, tt, 1.0));
            cs[k] = c;
            sn[k] = tt * c;
            on = 1;
            rotated = 1;
          }
        }
        pp[k] = p; qq[k] = q; act[k] = on;
      }
      __syncthreads();
      for (int w = threadIdx.x; w < half * R; w += blockDim.x) {       // rows p, q
        int k = w / R, j = w % R;
        if (act[k]) {
          int p = pp[k], q = qq[k];
          double c = cs[k], s = sn[k];
          double x = a[p * R + j], y = a[q * R + j];
          a[p * R + j] = fma(c, x, -(s * y));
          a[q * R + j] = fma(s, x, c * y);
        }
      }
      __syncthreads();
      for (int w = threadIdx.x; w < half * R; w += blockDim.x) {       // columns p, q
        int k = w / R, i = w % R;
        if (act[k]) {
          int p = pp[k], q = qq[k];
          double c = cs[k], s = sn[k];
          double x = a[i * R + p], y = a[i * R + q];
          a[i * R + p] = fma(c, x, -(s * y));
          a[i * R + q] = fma(s, x, c * y);
        }
      }
      __syncthreads();
      if (threadIdx.x < half && act[threadIdx.x]) {
        int p = pp[threadIdx.x], q = qq[threadIdx.x];
        a[p * R + q] = 0.0;
        a[q * R + p] = 0.0;
      }
      __syncthreads();
    }
    if (!rotated) break;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double mx = fabs(a[0]), mn = fabs(a[0]);
    for (int i = 1; i < R; ++i) {
      double d = fabs(a[i * R + i]);
      if (d > mx) mx = d;
      if (d < mn) mn = d;
    }
    // GLambda, NesterovMNLS.hpp:75-91
    double L = mx, mu = mn, lambda;
    double q = mu / L;
    if (1.0 / q > 1e6) lambda = 10.0 * mu;
    else if (1.0 / q > 1e3) lambda = mu;
    else lambda = mu / 10.0;
    L += lambda;
    mu += lambda;
    q = mu / L;
    prm->L0 = mx;
    prm->mu0 = mn;
    prm->L = L;
    prm->lambda = lambda;
    prm->q = q;
  }
}

__device__ __forceinline__ double update_alpha(double alpha, double q) {
  double a = 1.0;
  double b = alpha * alpha - q;
  double c = -alpha * alpha;
  double D = b * b - 4.0 * a * c;
  return (-b + sqrt(D)) / 2.0;
}

// One 64-lane wave per factor row; lane r holds column r.  Zt = V + lambda I lives in
// LDS; y_k is broadcast with readlane.  Arithmetic mirrors NesterovMNLS.hpp:205-231.
struct RowState {
  double y, acur, wt;
};

__device__ __forceinline__ double row_grad(const double *zt, int R, double y, double wt, int r) {
  double s = 0.0;
  for (int k = 0; k < R; ++k) s = fma(bcast_lane(y, k), zt[k * kMaxRank + r], s);
  return s - wt;
}

__global__ __launch_bounds__(256) void k_nnls_pass1(int64_t I, int R, int ldr, const double *__restrict__ V,
                                                    const NnlsParams *__restrict__ prm, const double *__restrict__ M,
                                                    const double *__restrict__ A, double d1, double d2, int max_inner,
                                                    unsigned long long *mask) {
  __shared__ double zt[kMaxRank * kMaxRank];
  const double lambda = prm->lambda, L = prm->L, q = prm->q;
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
    int a = e / R, b = e % R;
    zt[a * kMaxRank + b] = a == b ? V[e] + lambda : V[e];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= I) return;
  const bool on = lane < R;
  const int r = on ? lane : 0;
  double a0 = on ? A[row * ldr + r] : 0.0;
  double wt = on ? M[row * ldr + r] + lambda * a0 : 0.0;
  double y = a0, acur = a0, alpha = 1.0;
  unsigned long long bits[4] = {0ull, 0ull, 0ull, 0ull};
  for (int k = 0; k < max_inner; ++k) {
    double g = row_grad(zt, R, y, wt, r);
    double gy = on ? fabs(g * y) : 0.0;
    double gm = on ? g : INFINITY;
    double rmax = wave_max(gy), rmin = wave_min(gm);
    if (rmax <= d1 && rmin >= -d2) bits[k >> 6] |= 1ull << (k & 63);
    if (k == max_inner - 1) break;
    double tq = y - g / L;
    double an = tq > 0.0 ? tq : 0.0;
    double na = update_alpha(alpha, q);
    double beta = alpha * (1.0 - alpha) / (alpha * alpha + na);
    y = (1.0 + beta) * an - beta * acur;
    acur = an;
    alpha = na;
  }
  if (lane == 0) {
    for (int w = 0; w < 4; ++w) {
      unsigned long long m = bits[w];
      int lo = w * 64;
      if (max_inner < lo + 64) m |= (max_inner <= lo) ? ~0ull : (~0ull << (max_inner - lo));
      atomicAnd(&mask[w], m);
    }
  }
}

__device__ __forceinline__ int first_common_check(const unsigned long long *mask, int max_inner) {
  for (int w = 0; w < 4; ++w) {
    unsigned long long m = mask[w];
    if (m) {
      int k = w * 64 + __builtin_ctzll(m);
      return k < max_inner ? k : max_inner;
    }
  }
  return max_inner;
}

__global__ __launch_bounds__(256) void k_nnls_pass2(int64_t I, int R, int ldr, const double *__restrict__ V,
                                                    NnlsParams *prm, const double *__restrict__ M,
                                                    const double *__restrict__ Aold, double *__restrict__ Anew,
                                                    int max_inner, const unsigned long long *__restrict__ mask,
                                                    double *colsum) {
  __shared__ double zt[kMaxRank * kMaxRank];
  const double lambda = prm->lambda, L = prm->L, q = prm->q;
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
    int a = e / R, b = e % R;
    zt[a * kMaxRank + b] = a == b ? V[e] + lambda : V[e];
  }
  __syncthreads();
  const int kstar = first_common_check(mask, max_inner);
  if (blockIdx.x == 0 && threadIdx.x == 0) prm->iters = kstar;
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= I) return;
  const bool on = lane < R;
  const int r = on ? lane : 0;
  double a0 = on ? Aold[row * ldr + r] : 0.0;
  double wt = on ? M[row * ldr + r] + lambda * a0 : 0.0;
  double y = a0, acur = a0, alpha = 1.0;
  for (int k = 0; k < kstar; ++k) {
    double g = row_grad(zt, R, y, wt, r);
    double tq = y - g / L;
    double an = tq > 0.0 ? tq : 0.0;
    double na = update_alpha(alpha, q);
    double beta = alpha * (1.0 - alpha) / (alpha * alpha + na);
    y = (1.0 + beta) * an - beta * acur;
    acur = an;
    alpha = na;
  }
  if (lane < ldr) Anew[row * ldr + lane] = on ? acur : 0.0;
  if (on) atomicAdd(&colsum[r], fabs(acur));
}

__global__ void k_init_mask(unsigned long long *mask, double *colsum, int R) {
  if (threadIdx.x < 4) mask[threadIdx.x] = ~0ull;
  if ((int)threadIdx.x < R) colsum[threadIdx.x] = 0.0;
}

// min column abs-sum == 0 -> A = 0.9 A + 0.1 A_old (Cpd.hpp:206-207)
__global__ void k_zero_col_fallback(int64_t I, int R, int ldr, const double *colsum, const double *Aold, double *A,
                                    NnlsParams *prm) {
  __shared__ int hit;
  if (threadIdx.x == 0) {
    double mn = INFINITY;
    for (int r = 0; r < R; ++r) mn = colsum[r] < mn ? colsum[r] : mn;
    hit = mn == 0.0;
    prm->fallback = hit;
  }
  __syncthreads();
  if (!hit) return;
  for (int64_t e = threadIdx.x; e < I * ldr; e += blockDim.x) {
    if ((int)(e % ldr) < R) A[e] = 0.9 * A[e] + 0.1 * Aold[e];
  }
}

__global__ void k_dot_partials(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                               int64_t chunk, double *__restrict__ partials) {
  __shared__ double red[256];
  const int64_t e0 = (int64_t)blockIdx.x * chunk;
  const int64_t e1 = e0 + chunk < n ? e0 + chunk : n;
  double s = 0.0;
  for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) s = fma(a[e], b[e], s);
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < (int)blockDim.x; ++i) t += red[i];
    partials[blockIdx.x] = t;
  }
}

__device__ double vg_sum(int R, const double *V, const double *G) {
  double s = 0.0;
  for (int e = 0; e < R * R; ++e) s = fma(V[e], G[e], s);
  return s;
}

__global__ void k_cost_gram_final(int R, const double *frob, const double *ma_partials, int nparts, const double *V,
                                  const double *G, double *f, int accel) {
  if (threadIdx.x != 0) return;
  double ma = 0.0;
  for (int p = 0; p < nparts; ++p) ma += ma_partials[p];
  double vg = vg_sum(R, V, G);
  *f = accel ? sqrt(*frob + vg - 2.0 * ma) : sqrt(*frob - 2.0 * ma + vg);
}

// Normalize (Normalize.hpp:92-174): factors 0..N-2 other than the weight factor wf are
// scaled to unit columns, wf absorbs the product of their norms, the last factor is
// never touched and the zero test reads only the last lambda_fac entry; the Grams are
// rescaled by quotients.
struct NormArgs {
  int N, R, ldr, wf;
  int64_t dims[kMaxOrder];
  double *G[kMaxOrder];
  double *A[kMaxOrder];
};

__global__ __launch_bounds__(256) void k_normalize(NormArgs na, double *scratch) {
  const int N = na.N, R = na.R, last = N - 1;
  double *lam = scratch;                       // (N-1) x R
  double *scale = scratch + kMaxOrder * kMaxRank;  // N x R column scales (1 = untouched)
  __shared__ int nonzero[kMaxRank];
  if (threadIdx.x < R) {
    const int i = threadIdx.x;
    int pass = 0;
    for (int k = 0; k < N; ++k) {
      if (k != na.wf) lam[(k - pass) * R + i] = na.G[k][i * R + i];
      else pass = 1;
      scale[k * R + i] = 1.0;
    }
    double nf[kMaxOrder];
    int nz = 1;
    for (int j = 0; j < last; ++j) {
      nf[j] = sqrt(lam[j * R + i]);
      nz = !(nf[j] == 0.0);
    }
    nonzero[i] = nz;
    if (!nz) {
      for (int j = 0; j < last; ++j) lam[j * R + i] = 1.0;
    } else {
      double cumul = 1.0;
      pass = 0;
      for (int j = 0; j < last; ++j) {
        if (j == na.wf) { pass = 1; continue; }
        scale[j * R + i] = 1.0 / nf[j - pass];
        cumul *= nf[j - pass];
        lam[(j - pass) * R + i] = nf[j - pass];
      }
      scale[na.wf * R + i] = cumul;
    }
  }
  __syncthreads();
  for (int k = 0; k < N; ++k) {
    if (k == last && k != na.wf) continue;
    const int64_t n = na.dims[k] * na.ldr;
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
      int c = (int)(e % na.ldr);
      if (c < R && nonzero[c]) na.A[k][e] = na.A[k][e] * scale[k * R + c];
    }
  }
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
    int a = e / R, b = e % R;
    double nmw = 1.0;
    int pass = 0;
    for (int i = 0; i < last; ++i) {
      if (i == na.wf) { pass = 1; continue; }
      const double *l = lam + (i - pass) * R;
      double nm = l[a] * l[b];
      nmw = nmw * nm;
      na.G[i][e] = na.G[i][e] / nm;
    }
    na.G[na.wf][e] = na.G[na.wf][e] * nmw;
  }
}

__global__ void k_axpy_accel(int64_t n, double step, const double *old, const double *cur, double *out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    out[e] = old[e] + step * (cur[e] - old[e]);
}

int grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

void hadamard(hipStream_t s, int R, int ng, const double *const *grams, double *V) {
  GramPtrs G{};
  for (int j = 0; j < ng; ++j) G.p[j] = grams[j];
  k_hadamard<<<1, 256, 0, s>>>(R, ng, G, V);
}

void gram_partials(hipStream_t s, int64_t I, int R, int ldr, const double *A, double *partials, int nparts) {
  int64_t rchunk = (I + nparts - 1) / nparts;
  k_gram_partials<<<nparts, 256, 0, s>>>(I, R, ldr, A, rchunk, partials);
}

void gram_reduce(hipStream_t s, int R, const double *partials, int nparts, double *G) {
  k_gram_reduce<<<1, 256, 0, s>>>(R, partials, nparts, G);
}

void eig_extremes(hipStream_t s, int R, const double *V, NnlsParams *p) {
  k_eig_extremes<<<1, 256, 0, s>>>(R, V, p);
}

void nnls_init_mask(hipStream_t s, uint64_t *mask, double *colsum, int R) {
  k_init_mask<<<1, 64, 0, s>>>((unsigned long long *)mask, colsum, R);
}

void nnls_pass1(hipStream_t s, int64_t I, int R, int ldr, const double *V, const NnlsParams *p, const double *M,
                const double *A, double d1, double d2, int max_inner, uint64_t *mask) {
  k_nnls_pass1<<<(unsigned)((I + 3) / 4), 256, 0, s>>>(I, R, ldr, V, p, M, A, d1, d2, max_inner,
                                                       (unsigned long long *)mask);
}

void nnls_pass2(hipStream_t s, int64_t I, int R, int ldr, const double *V, NnlsParams *p, const double *M,
                const double *Aold, double *Anew, int max_inner, const uint64_t *mask, double *colsum) {
  k_nnls_pass2<<<(unsigned)((I + 3) / 4), 256, 0, s>>>(I, R, ldr, V, p, M, Aold, Anew, max_inner,
                                                       (const unsigned long long *)mask, colsum);
}

void zero_col_fallback(hipStream_t s, int64_t I, int R, int ldr, const double *colsum, const double *Aold, double *A,
                       NnlsParams *p) {
  k_zero_col_fallback<<<1, 256, 0, s>>>(I, R, ldr, colsum, Aold, A, p);
}

void dot_partials(hipStream_t s, int64_t n, const double *a, const double *b, double *partials, int nparts) {
  int64_t chunk = (n + nparts - 1) / nparts;
  k_dot_partials<<<nparts, 256, 0, s>>>(n, a, b, chunk, partials);
}

void cost_gram_final(hipStream_t s, int R, const double *frob, const double *ma_partials, int nparts, const double *V,
                     const double *G, double *f) {
  k_cost_gram_final<<<1, 64, 0, s>>>(R, frob, ma_partials, nparts, V, G, f, 0);
}

void accel_cost_final(hipStream_t s, int R, const double *frob, const double *ma_partials, int nparts, const double *V,
                      const double *G, double *f) {
  k_cost_gram_final<<<1, 64, 0, s>>>(R, frob, ma_partials, nparts, V, G, f, 1);
}

void normalize(hipStream_t s, int N, const int64_t *dims, int R, int ldr, int wf, double *const *grams,
               double *const *fac, double *scratch) {
  NormArgs na{};
  na.N = N; na.R = R; na.ldr = ldr; na.wf = wf;
  for (int k = 0; k < N; ++k) { na.dims[k] = dims[k]; na.G[k] = grams[k]; na.A[k] = fac[k]; }
  k_normalize<<<1, 256, 0, s>>>(na, scratch);
}

void axpy_accel(hipStream_t s, int64_t n, double step, const double *old, const double *cur, double *out) {
  k_axpy_accel<<<grid_for(n, 256), 256, 0, s>>>(n, step, old, cur, out);
}

}  // namespace ptn
