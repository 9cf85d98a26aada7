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
#include "wave_common.inc"

struct MutPtrs {
  double *p[kMaxOrder];
};

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
__device__ __forceinline__ void gram_rows32(const double (*As)[kMaxRank + 4], int R, int tid, double *o) {
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
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
      if (a0 + x < R && b0 + y < R) o[(a0 + x) * R + b0 + y] = acc[x][y];
}

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
  gram_rows32(As, R, tid, partials + (int64_t)blockIdx.x * R * R);
}

__global__ __launch_bounds__(256) void k_gram_combine(int R, int nchunks, const double *__restrict__ partials,
                                                      double *__restrict__ G) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < R * R; e += gridDim.x * blockDim.x) {
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

// With ng > 0 the matrix is first formed as the Hadamard product of the ng Grams in
// list order (V = G_l0, V = G_lj * V; PartialCwiseProd) and stored to Vout; with
// betas != nullptr the momentum coefficients of the Nesterov solve follow the
// eigenvalues (one kernel prepares an NNLS update).  One block of four waves.
template <int RB>
__global__ __launch_bounds__(256) void k_eig_extremes(int R, const double *__restrict__ V, GramPtrs G, int ng,
                                                      double *__restrict__ Vout, NnlsParams *prm, int max_inner,
                                                      double *__restrict__ betas) {
  __shared__ EigLds<RB> S;
  __shared__ double part[4][RB];
  const int n = R;
  for (int x = threadIdx.x; x < RB * RB; x += 256) {
    const int r = x / RB, c = x % RB;
    double v = 0.0;
    if (r < n && c < n) {
      const int e = r * n + c;
      if (ng > 0) {
        v = G.p[0][e];
        for (int j = 1; j < ng; ++j) v = G.p[j][e] * v;
        Vout[e] = v;
      } else {
        v = V[e];
      }
    }
    S.t[x] = v;
  }
  __syncthreads();
  double amax, amin;
  eig_extremes_block<RB>(n, S, part, amax, amin);
  if (threadIdx.x == 0) glambda_betas(amax, amin, prm, max_inner, betas);
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
                                                      const double *__restrict__ S, double d1, double d2, int max_inner,
                                                      unsigned long long *mask, double *__restrict__ hist,
                                                      double *colsum) {
  __shared__ double sbeta[kMaxInner + 1];
  if (blockIdx.x == 0 && (int)threadIdx.x < R) colsum[threadIdx.x] = 0.0;   // summed by k_nnls_select
  __shared__ double ybuf[4][RB];
  __shared__ unsigned long long wbits[4][4];
  for (int k = threadIdx.x; k < max_inner; k += blockDim.x) sbeta[k] = betas[k];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t row = (int64_t)blockIdx.x * 4 + wave;
  unsigned long long bits[4] = {~0ull, ~0ull, ~0ull, ~0ull};
  if (row < I) {
  const double lambda = prm->lambda, L = prm->L;
  const double invL = 1.0 / L;
  const bool on = lane < R;
  const int r = on ? lane : 0;
  double z[RB];
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    const double vk = (on && k < R) ? V[k * R + r] : 0.0;
    z[k] = (k == r && on) ? vk + lambda : vk;
  }
  // lanes >= R: zero start, zero column of Zt and zero Wt keep every iterate at +0.
  // The anchor of W~ is always the previous row A; the start is A or the warm start S.
  const double anc = on ? A[row * ldr + r] : 0.0;
  const double a0 = on ? (S ? S[row * ldr + r] : anc) : 0.0;
  const double wt = on ? M[row * ldr + r] + lambda * anc : 0.0;
  double y = a0, acur = a0;
  const int64_t hstride = I * ldr;
  double *hrow = hist + row * ldr + lane;
  if (lane < ldr) hrow[0] = a0;
  for (int kw = 0; kw < 4; ++kw) {
    unsigned long long m = 0ull;
    const int kend = max_inner - 64 * kw < 64 ? max_inner - 64 * kw : 64;
    for (int kb = 0; kb < kend; ++kb) {
      const int k = 64 * kw + kb;
      const double beta = sbeta[k];
      if (lane < RB) ybuf[wave][lane] = y;
      __builtin_amdgcn_wave_barrier();
      double s4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kk = 0; kk < RB; ++kk) s4[kk & 3] = fma(ybuf[wave][kk], z[kk], s4[kk & 3]);
      __builtin_amdgcn_wave_barrier();
      const double g = ((s4[0] + s4[1]) + (s4[2] + s4[3])) - wt;
      const bool fail = on && ((fabs(g * y) > d1) || (g < -d2));
      if (__ballot(fail) == 0ull) m |= 1ull << kb;       // the row passes check k
      // g / L as RN(g / L): reciprocal, then one exact-residual correction (Markstein)
      const double q0 = g * invL;
      const double tq = y - fma(fma(-q0, L, g), invL, q0);
      const double an = tq > 0.0 ? tq : 0.0;
      y = (1.0 + beta) * an - beta * acur;
      acur = an;
      if (lane < ldr) hrow[(k + 1) * hstride] = acur;
    }
    bits[kw] = m;
  }
  }
  // the block's mask (AND over its rows) goes to its own slot: no atomics
  if (lane == 0)
    for (int w = 0; w < 4; ++w) wbits[wave][w] = bits[w];
  __syncthreads();
  if (threadIdx.x < 4) {
    const int w = threadIdx.x;
    unsigned long long m = ((wbits[0][w] & wbits[1][w]) & wbits[2][w]) & wbits[3][w];
    const int lo = w * 64;
    if (max_inner < lo + 64) m |= (max_inner <= lo) ? ~0ull : (~0ull << (max_inner - lo));
    mask[(int64_t)blockIdx.x * 4 + w] = m;
  }
}

// Momentum coefficients beta_k of NesterovMNLS.hpp:222-223 (shared by every row).
__global__ void k_nnls_prep(const NnlsParams *prm, int max_inner, double *betas) {
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

__device__ __forceinline__ int first_common_check(const volatile unsigned long long *mask, int max_inner) {
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
  __shared__ unsigned long long mred[256];
  __shared__ int sk;
  {
    // AND of the per-block masks of k_nnls_iterate (4 words per block)
    const int64_t nblk = (I + 3) / 4;
    unsigned long long m = ~0ull;
    for (int64_t b = threadIdx.x >> 2; b < nblk; b += 64) m &= mask[b * 4 + (threadIdx.x & 3)];
    mred[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o >= 4; o >>= 1) {
      if ((int)threadIdx.x < o) mred[threadIdx.x] &= mred[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) sk = first_common_check(mred, max_inner);
    __syncthreads();
  }
  const int kstar = sk;
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

// the zero-column fallback's mix (Cpd.hpp:206-207)
__device__ __forceinline__ double zero_col_mix(double a, double aold) { return 0.9 * a + 0.1 * aold; }

// k_nnls_select over the rows [32 b, 32 b + 32) of block b, which also leaves the
// block's partial Gram of the selected rows (k_gram_chunks' association) in
// partials[b]; k_gram_finish applies the zero-column test and combines.
__global__ __launch_bounds__(256) void k_nnls_finish(int64_t I, int R, int ldr, const double *__restrict__ hist,
                                                     const unsigned long long *__restrict__ mask, int max_inner,
                                                     double *__restrict__ Anew, double *colsum,
                                                     double *__restrict__ partials, NnlsParams *prm) {
  __shared__ double As[32][kMaxRank + 4];
  __shared__ unsigned long long mred[256];
  __shared__ int sk;
  const int tid = threadIdx.x;
  {
    const int64_t nblk = (I + 3) / 4;
    unsigned long long m = ~0ull;
    for (int64_t b = tid >> 2; b < nblk; b += 64) m &= mask[b * 4 + (tid & 3)];
    mred[tid] = m;
    __syncthreads();
    for (int o = 128; o >= 4; o >>= 1) {
      if (tid < o) mred[tid] &= mred[tid + o];
      __syncthreads();
    }
    if (tid == 0) sk = first_common_check(mred, max_inner);
    __syncthreads();
  }
  const int kstar = sk;
  if (blockIdx.x == 0 && tid == 0) prm->iters = kstar;
  const double *src = hist + (int64_t)kstar * I * ldr;
  const int64_t i0 = (int64_t)blockIdx.x * 32;
  for (int x = tid; x < 32 * (kMaxRank + 4); x += 256) {
    const int i = x / (kMaxRank + 4), c = x % (kMaxRank + 4);
    double v = 0.0;
    if (i0 + i < I && c < ldr) {
      v = src[(i0 + i) * ldr + c];
      Anew[(i0 + i) * ldr + c] = v;
    }
    As[i][c] = c < R ? v : 0.0;
  }
  __syncthreads();
  if (tid < R) {
    double cs = 0.0;
    for (int i = 0; i < 32; ++i) cs += fabs(As[i][tid]);
    atomicAdd(&colsum[tid], cs);
  }
  gram_rows32(As, R, tid, partials + (int64_t)blockIdx.x * R * R);
}

// G = A^T A from the chunk partials (k_gram_combine's order), after the zero-column test
// of update_factor (Cpd.hpp:206-207): when some column abs-sum is zero, block 0 alone
// recomputes G from the mixed rows 0.9 A + 0.1 A_old in the same association and then
// stores the mixed rows.  prm->fallback records the test.
__global__ __launch_bounds__(256) void k_gram_finish(int64_t I, int R, int ldr, int nchunks,
                                                     const double *__restrict__ partials, double *__restrict__ G,
                                                     const double *colsum, const double *Aold, double *A,
                                                     NnlsParams *prm) {
  double mn = INFINITY;
  for (int r = 0; r < R; ++r) mn = colsum[r] < mn ? colsum[r] : mn;
  const bool hit = mn == 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) prm->fallback = hit;
  if (!hit) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < R * R; e += gridDim.x * blockDim.x) {
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
    return;
  }
  if (blockIdx.x != 0) return;
  for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
    const int a = e / R, b = e % R;
    double g = 0.0;
    for (int c = 0; c < nchunks; ++c) {
      double p = 0.0;
      for (int64_t i = 32 * (int64_t)c; i < I && i < 32 * (int64_t)(c + 1); ++i)
        p = fma(zero_col_mix(A[i * ldr + a], Aold[i * ldr + a]), zero_col_mix(A[i * ldr + b], Aold[i * ldr + b]), p);
      g = c == 0 ? p : g + p;
    }
    G[e] = g;
  }
  __syncthreads();                            // every entry of G read A before it changes
  const int c = threadIdx.x & 63;
  if (c >= R) return;
  for (int64_t row = threadIdx.x >> 6; row < I; row += blockDim.x >> 6) {
    const int64_t e = row * ldr + c;
    A[e] = zero_col_mix(A[e], Aold[e]);
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

// Column norms, scales and the Gram rescaling of Normalize (one block).  Columns whose
// norms vanish keep scale 1 on every factor, so the row pass leaves them bit-identical.
// After every thread of the block has stored its part of a readback to mapped host
// memory: one thread fences at system scope and stores the sequence word (vector store).
__device__ __forceinline__ void publish_seq(unsigned *dst, unsigned seq) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence_system();
    *(volatile unsigned *)dst = seq;
  }
}

__global__ __launch_bounds__(256) void k_normalize(NormArgs na, double *scratch, ExportList ex) {
  // the sweep's readback first (it does not depend on Normalize)
  for (int q = 0; q < ex.count; ++q)
    for (int e = threadIdx.x; e < ex.words[q]; e += blockDim.x) ex.dst[q][e] = ex.src[q][e];
  if (ex.seq_dst) publish_seq(ex.seq_dst, ex.seq);
  const int N = na.N, R = na.R, last = N - 1;
  double *lam = scratch;                       // (N-1) x R
  double *scale = scratch + kMaxOrder * kMaxRank;  // N x R column scales (1 = untouched)
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

// Row pass of Normalize: A_k[row][c] *= scale[k][c]; block (row tile of 4, factor k).
// With ax.out the line-search step of k_axpy_accel follows for every entry of the row
// (padding included): A_hat = A_old + step (A - A_old), A_old = A, snap = A.
__global__ __launch_bounds__(256) void k_normalize_rows(NormArgs na, const double *__restrict__ scratch,
                                                        AxpyArgs ax) {
  const int k = blockIdx.y, c = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= na.dims[k]) return;
  const bool scaled = !(k == na.N - 1 && k != na.wf) && c < na.R;
  double *a = na.A[k] + row * na.ldr + c;
  if (!ax.out) {
    if (scaled) *a = *a * scratch[kMaxOrder * kMaxRank + k * na.R + c];
    return;
  }
  if (c >= na.ldr) return;
  double v = *a;
  if (scaled) {
    v = v * scratch[kMaxOrder * kMaxRank + k * na.R + c];
    *a = v;
  }
  const double step = (ax.coeff == nullptr || *ax.coeff == ax.c) ? ax.step_c : ax.step_c1;
  const int64_t e = (na.A[k] - ax.cur) + row * na.ldr + c;
  const double o = ax.old[e];
  ax.out[e] = o + step * (v - o);
  ax.old[e] = v;
  if (ax.snap) ax.snap[e] = v;
}

// A_hat = A_old + step (A - A_old) (Cpd.hpp:251-258).  The step depends on the current
// acceleration coefficient, which the device updates: step_c for *coeff == c, else
// step_c1 (the coefficient rises by at most one per sweep).
// The same pass sets A_old = A for the next sweep (Cpd.hpp:348,353) and, when snap is
// given, copies A there too.
__global__ void k_axpy_accel(int64_t n, double step_c, double step_c1, int c, const int *coeff, double *old,
                             const double *cur, double *out, double *snap) {
  const double step = (coeff == nullptr || *coeff == c) ? step_c : step_c1;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const double o = old[e], a = cur[e];
    out[e] = o + step * (a - o);
    old[e] = a;
    if (snap) snap[e] = a;
  }
}

// Small device-to-host copies in one launch (4-byte words; dst is mapped host memory).
__global__ void k_export(ExportList x) {
  for (int q = 0; q < x.count; ++q)
    for (int e = threadIdx.x; e < x.words[q]; e += blockDim.x) x.dst[q][e] = x.src[q][e];
  if (x.seq_dst) publish_seq(x.seq_dst, x.seq);
}

// Line-search decision on the device (Cpd.hpp:262-269): accept the accelerated factors
// when f > f_hat, unless the run stops on the threshold at this sweep (the reference
// checks termination first, Cpd.hpp:339-346).  On acceptance the accelerated factors,
// their Grams and the mode-0 NNLS set-up prepared for them replace the current ones.
// flags[0] = accepted, flags[1] = threshold stop, flags[2] / flags[3] = accel_fail /
// accel_coeff, kept as line_search_accel keeps them (Cpd.hpp:270-277).
__global__ void k_accel_decide(const double *__restrict__ scal, double thr, int *flags, AccelCopy cp) {
  const double f = scal[1], fa = scal[2];
  const bool stop = f / sqrt(scal[0]) < thr;
  const bool accept = !stop && f > fa;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    flags[0] = accept ? 1 : 0;
    flags[1] = stop ? 1 : 0;
    if (!stop) {
      int fail = flags[2] + (accept ? 0 : 1);
      if (fail == 5) {
        fail = 0;
        flags[3] = flags[3] + 1;
      }
      flags[2] = fail;
    }
    if (cp.fa_out) {
      *cp.fa_out = fa;
      for (int q = 0; q < 4; ++q) cp.flags_out[q] = flags[q];
    }
    if (cp.f_out) *cp.f_out = f;
    if (cp.seq_dst) {                        // the sweep's readback is complete
      __threadfence_system();
      *(volatile unsigned *)cp.seq_dst = cp.seq;
    }
  }
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if (cp.pick_n) {
    const double *src = accept ? cp.pick_src[0] : cp.pick_src[1];
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < cp.pick_n; e += stride)
      cp.pick_dst[e] = src[e];
  }
  if (!accept) return;
  for (int q = 0; q < 4; ++q)
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < cp.n[q]; e += stride)
      cp.dst[q][e] = cp.src[q][e];
  if (cp.prm_src && blockIdx.x == 0 && threadIdx.x == 0) *cp.prm_dst = *cp.prm_src;
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
  k_gram_combine<<<(unsigned)((R * R + 63) / 64), 64, 0, s>>>(R, nchunks, partials, G);
}

void eig_extremes(hipStream_t s, int R, const double *V, NnlsParams *p) {
  const GramPtrs G{};
  if (R <= 16) k_eig_extremes<16><<<1, 256, 0, s>>>(R, V, G, 0, nullptr, p, 0, nullptr);
  else if (R <= 32) k_eig_extremes<32><<<1, 256, 0, s>>>(R, V, G, 0, nullptr, p, 0, nullptr);
  else k_eig_extremes<64><<<1, 256, 0, s>>>(R, V, G, 0, nullptr, p, 0, nullptr);
}

void nnls_setup(hipStream_t s, int R, int ng, const double *const *grams, double *V, NnlsParams *p, int max_inner,
                double *betas) {
  GramPtrs G{};
  for (int j = 0; j < ng; ++j) G.p[j] = grams[j];
  if (R <= 16) k_eig_extremes<16><<<1, 256, 0, s>>>(R, nullptr, G, ng, V, p, max_inner, betas);
  else if (R <= 32) k_eig_extremes<32><<<1, 256, 0, s>>>(R, nullptr, G, ng, V, p, max_inner, betas);
  else k_eig_extremes<64><<<1, 256, 0, s>>>(R, nullptr, G, ng, V, p, max_inner, betas);
}

void nnls_prep(hipStream_t s, const NnlsParams *p, int max_inner, double *betas) {
  k_nnls_prep<<<1, 64, 0, s>>>(p, max_inner, betas);
}

void nnls_iterate(hipStream_t s, int64_t I, int R, int ldr, const double *V, const NnlsParams *p,
                  const double *betas, const double *M, const double *A, const double *S, double d1, double d2,
                  int max_inner, uint64_t *mask, double *hist, double *colsum) {
  const unsigned grid = (unsigned)((I + 3) / 4);
  unsigned long long *m = (unsigned long long *)mask;
#define PTN_NI(RBV) \
  k_nnls_iterate<RBV><<<grid, 256, 0, s>>>(I, R, ldr, V, p, betas, M, A, S, d1, d2, max_inner, m, hist, colsum)
  if (R <= 16) PTN_NI(16);
  else if (R <= 32) PTN_NI(32);
  else PTN_NI(64);
#undef PTN_NI
}

void nnls_finish(hipStream_t s, int64_t I, int R, int ldr, const double *hist, const uint64_t *mask, int max_inner,
                 double *A, double *colsum, const double *Aold, double *G, double *partials, NnlsParams *p) {
  const int nchunks = (int)((I + 31) / 32);
  k_nnls_finish<<<(unsigned)nchunks, 256, 0, s>>>(I, R, ldr, hist, (const unsigned long long *)mask, max_inner, A,
                                                  colsum, partials, p);
  k_gram_finish<<<(unsigned)((R * R + 255) / 256), 256, 0, s>>>(I, R, ldr, nchunks, partials, G, colsum, Aold, A, p);
}

void nnls_select(hipStream_t s, int64_t I, int R, int ldr, const double *hist, const uint64_t *mask, int max_inner,
                 double *Anew, double *colsum, NnlsParams *p) {
  int64_t g = (I + 3) / 4;
  if (g > 256) g = 256;
  k_nnls_select<<<(unsigned)g, 256, 0, s>>>(I, R, ldr, hist, (const unsigned long long *)mask, max_inner, Anew,
                                            colsum, p);
}

void cost_gram(hipStream_t s, int64_t nma, const double *M, const double *A, int R, const double *V, const double *G,
               const double *frob, double *f, int accel) {
  k_cost_gram<<<1, 1024, 0, s>>>(nma, M, A, R, V, G, frob, f, accel);
}

void normalize(hipStream_t s, int N, const int64_t *dims, int R, int ldr, int wf, double *const *grams,
               double *const *fac, double *scratch, const ExportList *ex, const AxpyArgs *ax) {
  NormArgs na{};
  na.N = N; na.R = R; na.ldr = ldr; na.wf = wf;
  for (int k = 0; k < N; ++k) { na.dims[k] = dims[k]; na.G[k] = grams[k]; na.A[k] = fac[k]; }
  int64_t maxI = 0;
  for (int k = 0; k < N; ++k) maxI = dims[k] > maxI ? dims[k] : maxI;
  ExportList e0{};
  AxpyArgs a0{};
  k_normalize<<<1, 256, 0, s>>>(na, scratch, ex ? *ex : e0);
  k_normalize_rows<<<dim3((unsigned)((maxI + 3) / 4), (unsigned)N), 256, 0, s>>>(na, scratch, ax ? *ax : a0);
}

void accel_decide(hipStream_t s, const double *scal, double thr, int *flags, const AccelCopy &cp) {
  k_accel_decide<<<grid_for(cp.n[0] > 4096 ? cp.n[0] : 4096, 256), 256, 0, s>>>(scal, thr, flags, cp);
}

void axpy_accel(hipStream_t s, int64_t n, double step_c, double step_c1, int c, const int *coeff, double *old,
                const double *cur, double *out, double *snap) {
  k_axpy_accel<<<grid_for(n, 256), 256, 0, s>>>(n, step_c, step_c1, c, coeff, old, cur, out, snap);
}

void export_words(hipStream_t s, const ExportList &x) { k_export<<<1, 256, 0, s>>>(x); }

}  // namespace ptn
