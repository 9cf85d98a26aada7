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

// G = A^T A: block c sums rows [32c, 32c+32) in order for every entry (4 x 4 entries
// per thread, rows staged in LDS, missing rows padded with zeros, which add exact
// zeros); k_gram_combine adds the chunk sums in order.  Same association as the CPU
// restatement.
__global__ __launch_bounds__(256) void k_gram_chunks(int64_t I, int R, int ldr, const double *__restrict__ A,
                                                     double *__restrict__ partials) {
  __shared__ double As[32][kMaxRank + 4];
  const int tid = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * 32;
  for (int x = tid; x < 32 * (kMaxRank + 4); x += 256) {
    const int i = x / (kMaxRank + 4), c = x % (kMaxRank + 4);
    As[i][c] = (i0 + i < I && c < R) ? A[(i0 + i) * ldr + c] : 0.0;
  }
  __syncthreads();
  const int nb = (R + 3) / 4;                 // 4-wide blocks per dimension
  if (tid >= nb * nb) return;
  const int a0 = 4 * (tid / nb), b0 = 4 * (tid % nb);
  double acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = 0.0;
#pragma unroll 8
  for (int i = 0; i < 32; ++i) {
    double va[4], vb[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) { va[x] = As[i][a0 + x]; vb[x] = As[i][b0 + x]; }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y) acc[x][y] = fma(va[x], vb[y], acc[x][y]);
  }
  double *o = partials + (int64_t)blockIdx.x * R * R;
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
      if (a0 + x < R && b0 + y < R) o[(a0 + x) * R + b0 + y] = acc[x][y];
}

__global__ __launch_bounds__(256) void k_gram_combine(int R, int nchunks, const double *__restrict__ partials,
                                                      double *__restrict__ G) {
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
    double g = partials[e];
    int c = 1;
    for (; c + 4 <= nchunks; c += 4) {
      const double p0 = partials[(int64_t)c * R * R + e], p1 = partials[(int64_t)(c + 1) * R * R + e];
      const double p2 = partials[(int64_t)(c + 2) * R * R + e], p3 = partials[(int64_t)(c + 3) * R * R + e];
      g = (((g + p0) + p1) + p2) + p3;
    }
    for (; c < nchunks; ++c) g = g + partials[(int64_t)c * R * R + e];
    G[e] = g;
  }
}

constexpr double kPivminBase = 2.2250738585072014e-308;

__device__ int sturm_count(int n, const double *d, const double *e2, double pivmin, double x) {
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

// Multisection for the j-th smallest eigenvalue, run by one 32-lane half of wave 0:
// lane m < 31 probes lo + (m+1)(hi-lo)/32 each step; the new bracket is the first
// probe interval whose count exceeds j (the oracle scans the probes in order).
__device__ double sturm_find_half(int n, const double *d, const double *e2, double pivmin, double lo, double hi,
                                  int j, int m, int half) {
  for (int it = 0; it < 64; ++it) {
    const double w = hi - lo;
    const double sc = fabs(lo) > fabs(hi) ? fabs(lo) : fabs(hi);
    const bool stop = w <= 2.0 * kEps * sc || w <= pivmin;
    bool hit = false;
    if (!stop && m < 31) hit = sturm_count(n, d, e2, pivmin, fma(w, (double)(m + 1) * 0.03125, lo)) > j;
    const unsigned long long bal = __ballot(hit);
    const unsigned bits = (unsigned)(bal >> (32 * half)) & 0x7fffffffu;
    if (stop) break;
    const int ms = bits ? __builtin_ctz(bits) : 31;
    const double nlo = ms > 0 ? fma(w, (double)ms * 0.03125, lo) : lo;
    const double nhi = ms < 31 ? fma(w, (double)(ms + 1) * 0.03125, lo) : hi;
    if (nlo == lo && nhi == hi) break;
    lo = nlo;
    hi = nhi;
  }
  return 0.5 * (lo + hi);
}

// Extreme singular values (= max / min |eigenvalue|) of the symmetric V: Householder
// tridiagonalization (Golub & Van Loan Alg. 8.3.1) in LDS, then Sturm multisection
// (Alg. 8.4.1) on wave 0, the two halves bracketing two eigenvalues at once.  Sums
// run in index order with fma, matching the CPU restatement operation by operation;
// loops are unrolled to RB with zero padding (an added fma(x, 0, s) leaves s unchanged,
// s never being -0 here) so their LDS loads issue together.
template <int RB>
__global__ __launch_bounds__(256) void k_eig_extremes(int R, const double *__restrict__ V, NnlsParams *prm) {
  __shared__ double a[RB * RB];
  __shared__ double v[RB], p[RB], d[RB], e[RB], e2[RB];
  __shared__ double sh_beta, sh_K, sh_piv, sh_glo, sh_ghi, sh_res[4];
  __shared__ int sh_skip, sh_k0;
  const int n = R, tid = threadIdx.x;
  for (int x = tid; x < RB * RB; x += blockDim.x) {
    const int i = x / RB, j = x % RB;
    a[x] = (i < n && j < n) ? V[i * n + j] : 0.0;
  }
  if (tid < RB) { e[tid] = 0.0; v[tid] = 0.0; p[tid] = 0.0; }
  __syncthreads();
  for (int k = 0; k + 2 < n; ++k) {
    if (tid == 0) {
      double sigma = 0.0;
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        const double x = (i >= k + 2 && i < n) ? a[i * RB + k] : 0.0;
        sigma = fma(x, x, sigma);
      }
      const double x0 = a[(k + 1) * RB + k];
      if (sigma == 0.0) {
        e[k] = x0;
        sh_skip = 1;
      } else {
        const double mu = sqrt(fma(x0, x0, sigma));
        const double v0 = x0 <= 0.0 ? x0 - mu : x0 + mu;
        e[k] = x0 <= 0.0 ? mu : -mu;
        sh_beta = 2.0 / fma(v0, v0, sigma);
        sh_skip = 0;
        v[k + 1] = v0;
      }
    } else if (tid >= k + 2 && tid < n) {
      v[tid] = a[tid * RB + k];
    }
    if (tid == k) v[k] = 0.0;                 // entries <= k stay zero from here on
    __syncthreads();
    if (sh_skip) continue;                    // block-uniform
    const double beta = sh_beta;
    if (tid > k && tid < n) {
      double sacc = 0.0;
#pragma unroll
      for (int j = 0; j < RB; ++j) sacc = fma(a[tid * RB + j], v[j], sacc);
      p[tid] = beta * sacc;
    }
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
#pragma unroll
      for (int i = 0; i < RB; ++i) t = fma(p[i], v[i], t);
      sh_K = (0.5 * beta) * t;
    }
    __syncthreads();
    if (tid > k && tid < n) p[tid] = fma(-sh_K, v[tid], p[tid]);
    __syncthreads();
    const int m = n - k - 1;
    for (int x = tid; x < m * m; x += blockDim.x) {
      const int i = k + 1 + x / m, j = k + 1 + x % m;
      a[i * RB + j] = a[i * RB + j] - (v[i] * p[j] + p[i] * v[j]);
    }
    __syncthreads();
  }
  if (tid < n) d[tid] = a[tid * RB + tid];
  if (tid == 0 && n >= 2) e[n - 2] = a[(n - 1) * RB + (n - 2)];
  __syncthreads();
  if (tid == 0) {
    double emax = 1.0;
    for (int i = 0; i + 1 < n; ++i) {
      e2[i] = e[i] * e[i];
      if (e2[i] > emax) emax = e2[i];
    }
    const double pivmin = kPivminBase * emax;
    double glo = d[0], ghi = d[0];
    for (int i = 0; i < n; ++i) {
      const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0);
      const double lo = d[i] - r, hi = d[i] + r;
      if (lo < glo) glo = lo;
      if (hi > ghi) ghi = hi;
    }
    const double pad = 2.0 * kEps * (fabs(glo) > fabs(ghi) ? fabs(glo) : fabs(ghi)) + pivmin;
    sh_piv = pivmin;
    sh_glo = glo - pad;
    sh_ghi = ghi + pad;
    sh_k0 = sturm_count(n, d, e2, pivmin, 0.0);
  }
  __syncthreads();
  if (tid < 64) {                                          // wave 0
    const int half = tid >> 5, mm = tid & 31;
    const int k0 = sh_k0;
    const int j1 = half == 0 ? 0 : n - 1;
    const double r1 = sturm_find_half(n, d, e2, sh_piv, sh_glo, sh_ghi, j1, mm, half);
    double r2 = 0.0;
    const bool inner = k0 > 0 && k0 < n;                   // wave-uniform
    if (inner) r2 = sturm_find_half(n, d, e2, sh_piv, sh_glo, sh_ghi, half == 0 ? k0 - 1 : k0, mm, half);
    if (mm == 0) {
      sh_res[half] = r1;
      sh_res[2 + half] = r2;
    }
  }
  __syncthreads();
  if (tid == 0) {
    const double lmin = sh_res[0], lmax = sh_res[1];
    const int k0 = sh_k0;
    double amin;
    if (k0 == 0) amin = fabs(lmin);
    else if (k0 == n) amin = fabs(lmax);
    else {
      const double a1 = fabs(sh_res[2]), a2 = fabs(sh_res[3]);
      amin = a1 < a2 ? a1 : a2;
    }
    const double amax = fabs(lmin) > fabs(lmax) ? fabs(lmin) : fabs(lmax);
    // GLambda, NesterovMNLS.hpp:75-91
    double L = amax, mu = amin, lambda;
    double q = mu / L;
    if (1.0 / q > 1e6) lambda = 10.0 * mu;
    else if (1.0 / q > 1e3) lambda = mu;
    else lambda = mu / 10.0;
    L += lambda;
    mu += lambda;
    q = mu / L;
    prm->L0 = amax;
    prm->mu0 = amin;
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

// One 64-lane wave per factor row; lane r holds column r and keeps column r of
// Zt = V + lambda I in registers.  Each step the row y is broadcast through LDS and
// (Y Zt)(i, r) accumulates four interleaved partial sums over k mod 4, combined
// pairwise, as the CPU restatement does; lanes and k beyond R hold exact zeros, which
// leave the sums unchanged.  Each lane records whether its entry passes the stop test
// (NesterovMNLS.hpp:217) at every check; the row's mask is the AND over its lanes
// (max |G.Y| <= d1 and min G >= -d2 hold iff they hold entry-wise).  Iterate A_k goes
// to hist[k]; the common stopping point is selected afterwards.  The momentum
// coefficients depend only on q and come precomputed (k_nnls_prep).
template <int RB>
__global__ __launch_bounds__(256) void k_nnls_iterate(int64_t I, int R, int ldr, const double *__restrict__ V,
                                                      const NnlsParams *__restrict__ prm,
                                                      const double *__restrict__ betas,
                                                      const double *__restrict__ M, const double *__restrict__ A,
                                                      double d1, double d2, int max_inner,
                                                      unsigned long long *mask, double *__restrict__ hist) {
  __shared__ double sbeta[kMaxInner + 1];
  __shared__ double ybuf[4][RB];
  for (int k = threadIdx.x; k < max_inner; k += blockDim.x) sbeta[k] = betas[k];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t row = (int64_t)blockIdx.x * 4 + wave;
  if (row >= I) return;
  const double lambda = prm->lambda, L = prm->L;
  const bool on = lane < R;
  const int r = on ? lane : 0;
  double z[RB];
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    const double vk = (on && k < R) ? V[k * R + r] : 0.0;
    z[k] = (k == r && on) ? vk + lambda : vk;
  }
  const double a0 = on ? A[row * ldr + r] : 0.0;
  const double wt = on ? M[row * ldr + r] + lambda * a0 : 0.0;
  double y = a0, acur = a0;
  const int64_t hstride = I * ldr;
  double *hrow = hist + row * ldr + lane;
  if (lane < ldr) hrow[0] = a0;
  unsigned long long bits[4] = {0ull, 0ull, 0ull, 0ull};
  for (int k = 0; k < max_inner; ++k) {
    if (lane < RB) ybuf[wave][lane] = y;
    __builtin_amdgcn_wave_barrier();
    double s4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < RB; ++kk) s4[kk & 3] = fma(ybuf[wave][kk], z[kk], s4[kk & 3]);
    __builtin_amdgcn_wave_barrier();
    const double g = ((s4[0] + s4[1]) + (s4[2] + s4[3])) - wt;
    const bool pass = !on || (!(fabs(g * y) > d1) && !(g < -d2));
    if (pass) bits[k >> 6] |= 1ull << (k & 63);
    const double tq = y - g / L;
    const double an = tq > 0.0 ? tq : 0.0;
    const double beta = sbeta[k];
    y = on ? (1.0 + beta) * an - beta * acur : 0.0;
    acur = an;
    if (lane < ldr) hrow[(k + 1) * hstride] = on ? acur : 0.0;
  }
  for (int w = 0; w < 4; ++w) {
    unsigned long long m = bits[w];
    for (int o = 32; o > 0; o >>= 1) m &= __shfl_xor(m, o, 64);
    bits[w] = m;
  }
  if (lane == 0) {
    for (int w = 0; w < 4; ++w) {
      unsigned long long m = bits[w];
      const int lo = w * 64;
      if (max_inner < lo + 64) m |= (max_inner <= lo) ? ~0ull : (~0ull << (max_inner - lo));
      atomicAnd(&mask[w], m);
    }
  }
}

// Momentum coefficients beta_k of NesterovMNLS.hpp:222-223 (shared by every row), and
// the reset of the satisfaction mask and the column sums.
__global__ void k_nnls_prep(const NnlsParams *prm, int max_inner, double *betas, unsigned long long *mask,
                            double *colsum, int R) {
  if (threadIdx.x < 4) mask[threadIdx.x] = ~0ull;
  if ((int)threadIdx.x < R) colsum[threadIdx.x] = 0.0;
  if (threadIdx.x == 0) {
    const double q = prm->q;
    double alpha = 1.0;
    for (int k = 0; k < max_inner; ++k) {
      const double na = update_alpha(alpha, q);
      betas[k] = alpha * (1.0 - alpha) / (alpha * alpha + na);
      alpha = na;
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

// A_new = the iterate at the first check every row satisfies (else after max_inner
// steps); column abs-sums for the zero-column test (an exact-zero test, so the
// summation order does not matter).
__global__ __launch_bounds__(256) void k_nnls_select(int64_t I, int R, int ldr, const double *__restrict__ hist,
                                                     const unsigned long long *__restrict__ mask, int max_inner,
                                                     double *__restrict__ Anew, double *colsum, NnlsParams *prm) {
  __shared__ double red[4][kMaxRank];
  const int kstar = first_common_check(mask, max_inner);
  if (blockIdx.x == 0 && threadIdx.x == 0) prm->iters = kstar;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double *src = hist + (int64_t)kstar * I * ldr;
  double cs = 0.0;
  for (int64_t row = (int64_t)blockIdx.x * 4 + wave; row < I; row += (int64_t)gridDim.x * 4) {
    if (lane < ldr) {
      const double v = src[row * ldr + lane];
      Anew[row * ldr + lane] = v;
      cs += fabs(v);
    }
  }
  red[wave][lane] = cs;
  __syncthreads();
  if (wave == 0 && lane < R) atomicAdd(&colsum[lane], ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane]);
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
  const int c = threadIdx.x & 63;
  if (c >= R) return;
  for (int64_t row = threadIdx.x >> 6; row < I; row += blockDim.x >> 6) {
    const int64_t e = row * ldr + c;
    A[e] = 0.9 * A[e] + 0.1 * Aold[e];
  }
}

// <M, A> over I x ldr (padding is zero in both) and <V, G> over R x R: per-thread
// strided partial sums, then a fixed-shape tree in LDS (deterministic).  The Gram-form
// cost sqrt(||X||^2 - 2<M,A> + <V,G>) (CpdCUDA.hpp:137-142) or, for the accelerated
// factors, sqrt(||X||^2 + <V,G> - 2<M,A>) (Cpd.hpp:162-169).
__global__ __launch_bounds__(1024) void k_cost_gram(int64_t nma, const double *__restrict__ M,
                                                    const double *__restrict__ A, int R, const double *__restrict__ V,
                                                    const double *__restrict__ G, const double *frob, double *f,
                                                    int accel) {
  __shared__ double rma[1024], rvg[1024];
  const int tid = threadIdx.x;
  double ma = 0.0, vg = 0.0;
  for (int64_t e = tid; e < nma; e += 1024) ma = fma(M[e], A[e], ma);
  for (int e = tid; e < R * R; e += 1024) vg = fma(V[e], G[e], vg);
  rma[tid] = ma;
  rvg[tid] = vg;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (tid < w) {
      rma[tid] += rma[tid + w];
      rvg[tid] += rvg[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) *f = accel ? sqrt(*frob + rvg[0] - 2.0 * rma[0]) : sqrt(*frob - 2.0 * rma[0] + rvg[0]);
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
  {
    const int c = threadIdx.x & 63;
    for (int k = 0; k < N; ++k) {
      if ((k == last && k != na.wf) || c >= R || !nonzero[c]) continue;
      const double sc = scale[k * R + c];
      for (int64_t row = threadIdx.x >> 6; row < na.dims[k]; row += blockDim.x >> 6)
        na.A[k][row * na.ldr + c] = na.A[k][row * na.ldr + c] * sc;
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

void gram(hipStream_t s, int64_t I, int R, int ldr, const double *A, double *G, double *partials) {
  const int nchunks = (int)((I + 31) / 32);
  k_gram_chunks<<<nchunks, 256, 0, s>>>(I, R, ldr, A, partials);
  k_gram_combine<<<1, 256, 0, s>>>(R, nchunks, partials, G);
}

void eig_extremes(hipStream_t s, int R, const double *V, NnlsParams *p) {
  if (R <= 16) k_eig_extremes<16><<<1, 256, 0, s>>>(R, V, p);
  else if (R <= 32) k_eig_extremes<32><<<1, 256, 0, s>>>(R, V, p);
  else k_eig_extremes<64><<<1, 256, 0, s>>>(R, V, p);
}

void nnls_prep(hipStream_t s, const NnlsParams *p, int max_inner, double *betas, uint64_t *mask, double *colsum,
               int R) {
  k_nnls_prep<<<1, 64, 0, s>>>(p, max_inner, betas, (unsigned long long *)mask, colsum, R);
}

void nnls_iterate(hipStream_t s, int64_t I, int R, int ldr, const double *V, const NnlsParams *p,
                  const double *betas, const double *M, const double *A, double d1, double d2, int max_inner,
                  uint64_t *mask, double *hist) {
  const unsigned grid = (unsigned)((I + 3) / 4);
  unsigned long long *m = (unsigned long long *)mask;
  if (R <= 16) k_nnls_iterate<16><<<grid, 256, 0, s>>>(I, R, ldr, V, p, betas, M, A, d1, d2, max_inner, m, hist);
  else if (R <= 32) k_nnls_iterate<32><<<grid, 256, 0, s>>>(I, R, ldr, V, p, betas, M, A, d1, d2, max_inner, m, hist);
  else k_nnls_iterate<64><<<grid, 256, 0, s>>>(I, R, ldr, V, p, betas, M, A, d1, d2, max_inner, m, hist);
}

void nnls_select(hipStream_t s, int64_t I, int R, int ldr, const double *hist, const uint64_t *mask, int max_inner,
                 double *Anew, double *colsum, NnlsParams *p) {
  int64_t g = (I + 3) / 4;
  if (g > 256) g = 256;
  k_nnls_select<<<(unsigned)g, 256, 0, s>>>(I, R, ldr, hist, (const unsigned long long *)mask, max_inner, Anew,
                                            colsum, p);
}

void zero_col_fallback(hipStream_t s, int64_t I, int R, int ldr, const double *colsum, const double *Aold, double *A,
                       NnlsParams *p) {
  k_zero_col_fallback<<<1, 256, 0, s>>>(I, R, ldr, colsum, Aold, A, p);
}

void cost_gram(hipStream_t s, int64_t nma, const double *M, const double *A, int R, const double *V, const double *G,
               const double *frob, double *f, int accel) {
  k_cost_gram<<<1, 1024, 0, s>>>(nma, M, A, R, V, G, frob, f, accel);
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
