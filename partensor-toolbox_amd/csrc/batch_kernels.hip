// Batched independent non-negative CPD-ALS: many small order-3 problems of uneven sizes
// (each mode up to 128, rank up to 16) advanced one ALS sweep per launch of one
// persistent kernel.  A block takes whole problems from a per-sweep queue (largest
// first) and runs the reference's sweep on each (Cpd.hpp:288-356 with the per-mode
// MTTKRP, update_factor with NesterovMNLS, Gram, Gram-form cost, Normalize,
// line_search_accel), every step inside the block: the factors being multiplied, the
// MTTKRP and the Nesterov points sit in LDS, the tensor streams from HBM in FP32 and
// all arithmetic is FP64.  The sweep index is read from device memory, so one captured
// sweep (this kernel and the counter bump) replays for every sweep.
#include "ptn_internal.hpp"

namespace ptn {

namespace {

#include "wave_common.inc"

constexpr int kBR = 16;       // rank padding of a batched problem (rows of kBR doubles)
constexpr int kBMax = 128;    // largest mode size
constexpr int kPartial = 2 * kBMax * kBR;   // doubles available for the mode-0 group partials
constexpr int kLoads = 8;                   // tensor loads a lane issues before using them

struct BatchLds {
  double F[2][kBMax * kBR];   // the two factors a contraction multiplies by; during a solve
                              // the two Nesterov extrapolation points; F[0] then the result
  double M[kBMax * kBR];      // MTTKRP of the mode being updated
  double G[3][kBR * kBR];     // Grams of the current factors
  double V[kBR * kBR];        // Hadamard product for the mode being updated
  double Gh[3][kBR * kBR];    // Grams of the extrapolated factors
  EigLds<kBR> eig;
  NnlsParams prm;
  double betas[kMaxInner + 1];
  double red[BatchArgs::kMaxThreads];
  double lam[kBR];
  double fval;
  int colnz[kBR];
  int iters[3];
};

// threadIdx.x behind an empty asm statement: index arithmetic derived from it is redone
// in each routine instead of being hoisted out of the problem loop and held in registers
__device__ __forceinline__ int tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

template <int BS>
__device__ __forceinline__ double block_sum(double v, double *red) {
  // fixed-shape tree over the block (deterministic)
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int w = BS / 2; w > 0; w >>= 1) {
    if (t < w) red[t] = red[t] + red[t + w];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
  return v;
}

// rows [0, I) of a factor (global, stride kBR) into LDS
template <int BS>
__device__ __forceinline__ void load_rows(const double *src, int I, double *dst) {
  for (int e = threadIdx.x; e < I * kBR; e += BS) dst[e] = src[e];
}

// Gram G = A^T A for the first R columns, rows summed in order in 32-row chunks and the
// chunk sums added in order (the order of the CPU restatement's ptno_gram)
template <int BS>
__device__ __forceinline__ void gram_block(const double *A, int I, int R, double *G) {
  const int t = tid();
  if (t < kBR * kBR) {
    const int a = t / kBR, b = t % kBR;
    double g = 0.0;
    if (a < R && b < R) {
      for (int c0 = 0; c0 < I; c0 += 32) {
        double s = 0.0;
        const int c1 = c0 + 32 < I ? c0 + 32 : I;
        for (int i = c0; i < c1; ++i) s = fma(A[i * kBR + a], A[i * kBR + b], s);
        g = c0 == 0 ? s : g + s;
      }
    }
    G[t] = g;
  }
  __syncthreads();
}

// Mode-0 MTTKRP, first stage.  Thread (i0, g) of the first G * I0 threads takes the
// slabs i2 = g, g + G, ...: T = sum_i1 X(i0, i1, i2) F1[i1] (the loads of one i1 are
// contiguous over i0), acc += T * F2[i2].
template <int RP>
__device__ __forceinline__ void mttkrp0_partial(const float *X, int I0, int I1, int I2, int G, const double *F1,
                                                const double *F2, double (&acc)[RP]) {
  const int t = tid(), g = t / I0, i0 = t - g * I0;
#pragma unroll
  for (int r = 0; r < RP; ++r) acc[r] = 0.0;
  if (g >= G) return;
  const int s2 = I0 * I1;
  for (int i2 = g; i2 < I2; i2 += G) {
    const float *x = X + (i0 + s2 * i2);
    double T[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) T[r] = 0.0;
    int i1 = 0;
    for (; i1 + kLoads <= I1; i1 += kLoads) {     // kLoads tensor loads in flight
      float xs[kLoads];
#pragma unroll
      for (int u = 0; u < kLoads; ++u) xs[u] = x[I0 * (i1 + u)];
#pragma unroll
      for (int u = 0; u < kLoads; ++u) {
        const double xv = (double)xs[u];
        const double *f = F1 + (i1 + u) * kBR;
#pragma unroll
        for (int r = 0; r < RP; ++r) T[r] = fma(xv, f[r], T[r]);
      }
    }
    for (; i1 < I1; ++i1) {
      const double xv = (double)x[I0 * i1];
      const double *f = F1 + i1 * kBR;
#pragma unroll
      for (int r = 0; r < RP; ++r) T[r] = fma(xv, f[r], T[r]);
    }
    const double *f2 = F2 + i2 * kBR;
#pragma unroll
    for (int r = 0; r < RP; ++r) acc[r] = fma(T[r], f2[r], acc[r]);
  }
}

// Mode-0 MTTKRP, second stage: the G group partials go through LDS (P, laid out
// [r][g][i0]) and are summed in group order into M; columns RP.. of M are zeroed.
template <int BS, int RP>
__device__ __forceinline__ void mttkrp0_combine(int I0, int G, const double (&acc)[RP], double *P, double *M) {
  const int t = tid(), g = t / I0, i0 = t - g * I0;
  if (g < G) {
#pragma unroll
    for (int r = 0; r < RP; ++r) P[(r * G + g) * I0 + i0] = acc[r];
  }
  __syncthreads();
  for (int e = t; e < I0 * kBR; e += BS) {
    const int i = e / kBR, r = e - i * kBR;
    double s = 0.0;
    if (r < RP) {
      s = P[r * G * I0 + i];
      for (int q = 1; q < G; ++q) s = s + P[(r * G + q) * I0 + i];
    }
    M[e] = s;
  }
  __syncthreads();
}

// Modes 1 and 2: wave w < nw takes the output rows io = w, w + nw, ...  Its lanes split
// the row's slab into i0 (the contiguous index; a lane takes i0 and i0 + 64 when
// I0 > 64) and phases of the inner index j (when I0 < 64):
// T = sum_j X F_j[j], P = sum_i0 T * F0[i0], P summed over the wave.
// MODE 1: out i1, j = i2 (element i0 + I0 (io + I1 j)); MODE 2: out i2, j = i1.
template <int RP, int MODE>
__device__ __forceinline__ void mttkrp_rows(const float *X, int I0, int I1, int I2, const double *F0, const double *Fj, double *M,
                            int nw) {
  const int t = tid(), lane = t & 63, wave = t >> 6;
  if (wave >= nw) return;
  const int Iout = MODE == 1 ? I1 : I2, Iin = MODE == 1 ? I2 : I1;
  const int so = MODE == 1 ? I0 : I0 * I1;
  const int sj = MODE == 1 ? I0 * I1 : I0;
  int JP = 1, i0 = lane, ph = 0;
  if (I0 < 64) {
    JP = 64 / I0;
    ph = lane / I0;
    i0 = lane - ph * I0;
  }
  const bool live = ph < JP;
  for (int io = wave; io < Iout; io += nw) {
    double P[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) P[r] = 0.0;
    if (live) {
      for (int ia = i0; ia < I0; ia += 64) {
        double T[RP];
#pragma unroll
        for (int r = 0; r < RP; ++r) T[r] = 0.0;
        const float *x = X + (ia + so * io);
        int j = ph;
        for (; j + (kLoads - 1) * JP < Iin; j += kLoads * JP) {
          float xs[kLoads];
#pragma unroll
          for (int u = 0; u < kLoads; ++u) xs[u] = x[sj * (j + u * JP)];
#pragma unroll
          for (int u = 0; u < kLoads; ++u) {
            const double xv = (double)xs[u];
            const double *f = Fj + (j + u * JP) * kBR;
#pragma unroll
            for (int r = 0; r < RP; ++r) T[r] = fma(xv, f[r], T[r]);
          }
        }
        for (; j < Iin; j += JP) {
          const double xv = (double)x[sj * j];
          const double *f = Fj + j * kBR;
#pragma unroll
          for (int r = 0; r < RP; ++r) T[r] = fma(xv, f[r], T[r]);
        }
        const double *f0 = F0 + ia * kBR;
#pragma unroll
        for (int r = 0; r < RP; ++r) P[r] = fma(T[r], f0[r], P[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < RP; ++r) P[r] = wave_sum(P[r]);
    if (lane < kBR) {
      double v = 0.0;
#pragma unroll
      for (int r = 0; r < RP; ++r) v = lane == r ? P[r] : v;
      M[io * kBR + lane] = v;
    }
  }
}

// NesterovMNLS (NesterovMNLS.hpp:181-234) for the I x R entries of one factor.  Thread
// t = (row slot rs, column c), c = t mod R, takes rows rs, rs + NS, ... (NS = BS / R).
// Its entries of the iterate, of W~ = M + lambda A0 and of the next iterate stay in
// registers; the extrapolation point Y sits in LDS in two buffers, so one barrier per
// step both publishes the next point and ANDs the stopping test (the reference's global
// stop); the step after a satisfied test is computed and discarded.  Per entry the
// arithmetic is that of the single-problem kernel and the CPU restatement (four
// interleaved partial sums, g / L through the corrected reciprocal).  Writes the result,
// after the zero-column fallback (Cpd.hpp:206-207, an exact-zero test of each column),
// to Ag (global rows, which hold the start point on entry) and to S.F[0]; returns the
// step count.
template <int BS, int RP>
__device__ __forceinline__ int nnls_block(int I, int R, BatchLds &S, double *Ag, double d1, double d2, int max_inner) {
  constexpr int NJ = (kBMax + BS / RP - 1) / (BS / RP);   // rows per thread at the smallest NS
  const int NS = BS / R, t = tid(), rs = t / R, c = t - rs * R;
  const bool on = rs < NS;
  double *Yb0 = S.F[0], *Yb1 = S.F[1];
  const double lambda = S.prm.lambda, L = S.prm.L, invL = 1.0 / L;
  double z[RP];
#pragma unroll
  for (int k = 0; k < RP; ++k) {
    const double vk = k < R ? S.V[k * kBR + c] : 0.0;
    z[k] = k == c ? vk + lambda : vk;
  }
  double ac[NJ], an[NJ], wt[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int i = rs + NS * j;
    ac[j] = 0.0;
    an[j] = 0.0;
    wt[j] = 0.0;
    if (on && i < I) {
      const double a0 = Ag[i * kBR + c];
      ac[j] = a0;
      wt[j] = S.M[i * kBR + c] + lambda * a0;
      Yb0[i * kBR + c] = a0;
    }
  }
  for (int e = t; e < I * kBR; e += BS)
    if (e % kBR >= R) {
      Yb0[e] = 0.0;
      Yb1[e] = 0.0;
    }
  if (t < kBR) S.colnz[t] = 0;
  __syncthreads();
  int it = 0, b = 0;
  for (;;) {
    const double *Y = b ? Yb1 : Yb0;
    double *Yn = b ? Yb0 : Yb1;
    const double beta = it < max_inner ? S.betas[it] : 0.0;
    int pass = 1;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int i = rs + NS * j;
      if (on && i < I) {
        const double *yr = Y + i * kBR;
        double s4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < RP; ++k) s4[k & 3] = fma(yr[k], z[k], s4[k & 3]);
        const double g = ((s4[0] + s4[1]) + (s4[2] + s4[3])) - wt[j];
        const double y = yr[c];
        if ((fabs(g * y) > d1) || (g < -d2)) pass = 0;
        const double q0 = g * invL;
        const double tq = y - fma(fma(-q0, L, g), invL, q0);
        const double a = tq > 0.0 ? tq : 0.0;
        an[j] = a;
        Yn[i * kBR + c] = (1.0 + beta) * a - beta * ac[j];
      }
    }
    const int all = __syncthreads_and(pass);
    if (all || it >= max_inner) break;
#pragma unroll
    for (int j = 0; j < NJ; ++j) ac[j] = an[j];
    b ^= 1;
    ++it;
  }
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int i = rs + NS * j;
    if (on && i < I && ac[j] != 0.0) S.colnz[c] = 1;
  }
  __syncthreads();
  int fb = 0;
  for (int k = 0; k < R; ++k) fb |= S.colnz[k] == 0;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int i = rs + NS * j;
    if (on && i < I) {
      double v = ac[j];
      if (fb) v = 0.9 * v + 0.1 * Ag[i * kBR + c];
      Ag[i * kBR + c] = v;
      Yb0[i * kBR + c] = v;
    }
  }
  __syncthreads();
  return it;
}

// one sweep of one problem
template <int BS, int RP>
__device__ __forceinline__ void sweep_problem(const BatchArgs &a, int p, int sweep, BatchLds &S) {
  constexpr int NW = BS / 64, NT = BS - 64;    // the last wave solves the eigenproblem meanwhile
  const BatchDesc &d = a.desc[p];
  BatchProb &st = a.st[p];
  if (st.done) return;                                    // stopped on the threshold earlier
  const int I0 = d.I[0], I1 = d.I[1], I2 = d.I[2];
  const int R = d.R, t = threadIdx.x;
  const float *X = a.X + d.x_off;
  double *fac = a.fac + d.f_off * kBR;
  double *const A0 = fac, *const A1 = fac + (int64_t)d.rowoff[1] * kBR, *const A2 = fac + (int64_t)d.rowoff[2] * kBR;
  double *gram_g = a.grams + d.g_off;
  for (int e = t; e < 3 * kBR * kBR; e += BS) S.G[e / (kBR * kBR)][e % (kBR * kBR)] = gram_g[e];
  __syncthreads();
#pragma unroll 1
  for (int n = 0; n < 3; ++n) {
    // V_n: Hadamard of the other Grams (PartialCwiseProd); operands of the MTTKRP
    const int o1 = n == 0 ? 1 : 0, o2 = n == 2 ? 1 : 2;
    if (t < kBR * kBR) S.V[t] = S.G[o2][t] * S.G[o1][t];
    const int Ia = n == 0 ? I1 : I0, Ib = n == 2 ? I1 : I2;
    load_rows<BS>(n == 0 ? A1 : A0, Ia, S.F[0]);
    load_rows<BS>(n == 2 ? A1 : A2, Ib, S.F[1]);
    __syncthreads();
    const int G0 = (NT < kPartial / RP ? NT : kPartial / RP) / I0;
    double acc[RP];
    if (t >= NT) {
      const int l = t & 63;
      for (int x = l; x < kBR * kBR; x += 64) S.eig.t[x] = S.V[x];
      PTN_WAVE_SYNC();
      double amax, amin;
      eig_extremes_wave<kBR>(R, S.eig, amax, amin);
      if (l == 0) glambda_betas(amax, amin, &S.prm, a.max_inner, S.betas);
    } else if (n == 0) {
      mttkrp0_partial<RP>(X, I0, I1, I2, G0, S.F[0], S.F[1], acc);
    } else if (n == 1) {
      mttkrp_rows<RP, 1>(X, I0, I1, I2, S.F[0], S.F[1], S.M, NW - 1);
    } else {
      mttkrp_rows<RP, 2>(X, I0, I1, I2, S.F[0], S.F[1], S.M, NW - 1);
    }
    __syncthreads();
    if (n == 0) mttkrp0_combine<BS, RP>(I0, G0, acc, &S.F[0][0], S.M);
    const int In = n == 0 ? I0 : n == 1 ? I1 : I2;
    const int it = nnls_block<BS, RP>(In, R, S, n == 0 ? A0 : n == 1 ? A1 : A2, a.d1, a.d2, a.max_inner);
    if (t == 0) S.iters[n] = it;
    gram_block<BS>(S.F[0], In, R, S.G[n]);
    if (n == 2) {
      // Gram-form cost (CpdCUDA.hpp:137-142): sqrt(||X||^2 - 2 <M, A> + <V, G>)
      double ma = 0.0, vg = 0.0;
      for (int e = t; e < I2 * kBR; e += BS) ma = fma(S.M[e], S.F[0][e], ma);
      if (t < kBR * kBR) vg = S.V[t] * S.G[2][t];
      const double sma = block_sum<BS>(ma, S.red), svg = block_sum<BS>(vg, S.red);
      if (t == 0) S.fval = sqrt(d.frob - 2.0 * sma + svg);
      __syncthreads();
    }
  }
  // Normalize (Normalize.hpp:92-174) with weight factor 0: factor 1 to unit columns,
  // factor 0 takes the norms, factor 2 untouched; the zero test reads the last norm only
  if (a.normalization) {
    if (t < R) {
      const double n1 = sqrt(S.G[1][t * kBR + t]), n2 = sqrt(S.G[2][t * kBR + t]);
      S.lam[t] = n2 == 0.0 ? 1.0 : n1;
    }
    __syncthreads();
    for (int e = t; e < I0 * kBR; e += BS) {
      const int r = e % kBR;
      if (r < R) A0[e] = A0[e] * S.lam[r];
    }
    for (int e = t; e < I1 * kBR; e += BS) {
      const int r = e % kBR;
      if (r < R) A1[e] = A1[e] * (1.0 / S.lam[r]);
    }
    if (t < kBR * kBR) {
      const int ra = t / kBR, rb = t % kBR;
      if (ra < R && rb < R) {
        const double nm = S.lam[ra] * S.lam[rb];
        S.G[1][t] = S.G[1][t] / nm;
        S.G[0][t] = S.G[0][t] * nm;
      }
    }
    __syncthreads();
  }
  const double f = S.fval;
  const bool stop = f / sqrt(d.frob) < a.thr;
  double fa = 0.0;
  bool tried = false, accept = false;
  double *oldf = a.old + d.f_off * kBR, *accf = a.acc + d.f_off * kBR;
  if (!stop && sweep < a.max_iter && a.acceleration) {
    const int nrows = d.rowoff[2] + I2;
    if (sweep > 1) {
      // line_search_accel (Cpd.hpp:247-278): A^ = A_old + step (A - A_old), A_old = A
      tried = true;
      const double step = a.pow_table[(int64_t)sweep * a.pow_stride + st.accel_coeff];
      for (int e = t; e < nrows * kBR; e += BS) {
        const double cur = fac[e], o = oldf[e];
        accf[e] = o + step * (cur - o);
        oldf[e] = cur;
      }
      __syncthreads();
      const double *Ah0 = accf, *Ah1 = accf + (int64_t)d.rowoff[1] * kBR, *Ah2 = accf + (int64_t)d.rowoff[2] * kBR;
      load_rows<BS>(Ah2, I2, S.F[0]);
      __syncthreads();
      gram_block<BS>(S.F[0], I2, R, S.Gh[2]);
      load_rows<BS>(Ah0, I0, S.F[0]);
      load_rows<BS>(Ah1, I1, S.F[1]);
      __syncthreads();
      gram_block<BS>(S.F[0], I0, R, S.Gh[0]);
      gram_block<BS>(S.F[1], I1, R, S.Gh[1]);
      mttkrp_rows<RP, 2>(X, I0, I1, I2, S.F[0], S.F[1], S.M, NW);
      __syncthreads();
      double ma = 0.0, vg = 0.0;
      for (int e = t; e < I2 * kBR; e += BS) ma = fma(S.M[e], Ah2[e], ma);
      if (t < kBR * kBR) vg = (S.Gh[1][t] * S.Gh[0][t]) * S.Gh[2][t];
      const double sma = block_sum<BS>(ma, S.red), svg = block_sum<BS>(vg, S.red);
      fa = sqrt(d.frob + svg - 2.0 * sma);
      accept = f > fa;
      if (accept) {
        for (int e = t; e < nrows * kBR; e += BS) fac[e] = accf[e];
        for (int e = t; e < 3 * kBR * kBR; e += BS)
          S.G[e / (kBR * kBR)][e % (kBR * kBR)] = S.Gh[e / (kBR * kBR)][e % (kBR * kBR)];
      }
    } else {
      for (int e = t; e < nrows * kBR; e += BS) oldf[e] = fac[e];
    }
  }
  __syncthreads();
  for (int e = t; e < 3 * kBR * kBR; e += BS) gram_g[e] = S.G[e / (kBR * kBR)][e % (kBR * kBR)];
  if (t == 0) {
    st.f = accept ? fa : f;
    st.f_accel = fa;
    st.done = stop ? 1 : 0;
    st.sweeps = sweep;
    st.tried = tried ? 1 : 0;
    st.accepted = accept ? 1 : 0;
    for (int n = 0; n < 3; ++n) st.nnls_iters[n] = S.iters[n];
    if (tried && !accept) st.accel_fail += 1;
    if (st.accel_fail == 5) {
      st.accel_fail = 0;
      st.accel_coeff += 1;
    }
    if (a.log) {
      BatchLog &lg = a.log[(int64_t)p * a.max_iter + (sweep - 1)];
      lg.f_value = f;
      lg.f_accel = fa;
      for (int n = 0; n < 3; ++n) lg.nnls_iters[n] = S.iters[n];
      lg.accel_tried = tried ? 1 : 0;
      lg.accel_accepted = accept ? 1 : 0;
      lg.stopped = stop ? 1 : 0;
    }
  }
  __syncthreads();
}

template <int BS>
__global__ __launch_bounds__(BS, 2) void k_batch_sweep(BatchArgs a) {
  __shared__ BatchLds S;
  __shared__ int sp;
  const int sweep = *a.sweep_ctr;
  for (;;) {
    if (threadIdx.x == 0) sp = atomicAdd(&a.queue[sweep], 1);
    __syncthreads();
    const int q = sp;
    __syncthreads();
    if (q >= a.nprob) break;
    const int p = a.order[q];
    switch ((a.desc[p].R + 3) >> 2) {
      case 1: sweep_problem<BS, 4>(a, p, sweep, S); break;
      case 2: sweep_problem<BS, 8>(a, p, sweep, S); break;
      case 3: sweep_problem<BS, 12>(a, p, sweep, S); break;
      default: sweep_problem<BS, 16>(a, p, sweep, S); break;
    }
  }
}

__global__ void k_batch_next(int *ctr) {
  if (threadIdx.x == 0) *ctr = *ctr + 1;
}

// ||X_p||^2 and the Grams of the initial factors, one block per problem
template <int BS>
__global__ __launch_bounds__(BS) void k_batch_init(BatchArgs a) {
  __shared__ BatchLds S;
  const int p = blockIdx.x;
  BatchDesc &d = a.desc[p];
  const int64_t n = (int64_t)d.I[0] * d.I[1] * d.I[2];
  const float *X = a.X + d.x_off;
  double s = 0.0;
  for (int64_t e = threadIdx.x; e < n; e += BS) {
    const double x = (double)X[e];
    s = fma(x, x, s);
  }
  const double fro = block_sum<BS>(s, S.red);
  if (threadIdx.x == 0) d.frob = fro;
  const double *fac = a.fac + d.f_off * kBR;
  for (int k = 0; k < 3; ++k) {
    load_rows<BS>(fac + (int64_t)d.rowoff[k] * kBR, d.I[k], S.F[0]);
    __syncthreads();
    gram_block<BS>(S.F[0], d.I[k], d.R, S.G[k]);
    for (int e = threadIdx.x; e < kBR * kBR; e += BS) a.grams[d.g_off + k * kBR * kBR + e] = S.G[k][e];
    __syncthreads();
  }
}

}  // namespace

size_t batch_lds_bytes() { return sizeof(BatchLds); }

void batch_init(hipStream_t s, const BatchArgs &a) { k_batch_init<256><<<a.nprob, 256, 0, s>>>(a); }

void batch_sweep(hipStream_t s, const BatchArgs &a, int block, int grid) {
  if (block == 512) k_batch_sweep<512><<<grid, 512, 0, s>>>(a);
  else k_batch_sweep<256><<<grid, 256, 0, s>>>(a);
}

void batch_next_sweep(hipStream_t s, int *ctr) { k_batch_next<<<1, 64, 0, s>>>(ctr); }

}  // namespace ptn
