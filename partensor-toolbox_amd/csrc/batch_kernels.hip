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
  int seg[2 * BatchArgs::kMaxThreads / 64];   // partial MTTKRP tiles per wave
  uint64_t vote[2];           // Nesterov stop test: one byte per wave, two step parities
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

// The MTTKRP of one mode on v_mfma_f64_16x16x4f64:
//   M[row][r] = sum_k X(row, k) KR[k][r],  KR[k][r] = Fa[ia][r] Fb[ib][r],  k = ia + Ia ib,
//   MODE 0: row = i0, X offset row + I0 k,                 Fa = A1, Fb = A2
//   MODE 1: row = i1, X offset I0 row + ia + I0 I1 ib,      Fa = A0, Fb = A2
//   MODE 2: row = i2, X offset I0 I1 row + k,               Fa = A0, Fb = A1
// A wave works on a band of 32 rows (two 16 x 16 output tiles) in steps of 4 k: per step
// one Khatri-Rao fragment, two tensor fragments, two MFMAs.  Lane l holds A = X(row
// l % 16 of each tile, k = 4 s + l / 16) (FP32 widened to FP64), B = KR[4 s + l / 16][l % 16]
// (0 past K), and D rows l / 16 + 4 j, column l % 16.  Rows past the mode's size read a valid
// row and are not stored.  The bands x KS steps are split evenly over the nw waves
// (stream-K): a wave stores the bands its range covers whole straight into M and keeps the
// (at most two) partial bands in registers; mttkrp_mfma_partials / mttkrp_mfma_reduce add
// them up in wave order (a fixed order: the result does not depend on the grid).  LAST_Q < 4:
// the last wave takes LAST_Q quarters of a share (it has other work first), the others a share.
// F32: v_mfma_f32_16x16x4f32 on the FP32 tensor values and the FP32-rounded Khatri-Rao
// values (twice the FP64 matrix-core rate); each group of kSteps steps accumulates in FP32
// and is then added to the FP64 totals.  Its D layout is lane l, register j -> row
// 4 (l / 16) + j (the FP64 instruction's is row l / 16 + 4 j), both checked on gfx950.
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <bool F32>
__device__ __forceinline__ int dtile_row(int kk, int j) { return F32 ? 4 * kk + j : kk + 4 * j; }
constexpr int kSteps = 12;                  // k-steps whose tensor loads a lane issues together

// A raw buffer over a problem's tensor (bytes), its base in scalar registers: the group
// loads below take the lane's offset in one register and the step's in a scalar one, so
// the group's loads need no 64-bit address registers.  Reads past `bytes` return 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tensor_rsrc(const float *X, int bytes) {
  const uint64_t a = (uint64_t)X;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  return __builtin_amdgcn_make_buffer_rsrc((void *)(((uint64_t)hi << 32) | lo), (short)0,
                                           __builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}

struct MfmaPart {
  f64x4 a0, a1, b0, b1;                      // slot 0 (first partial band): tiles a0 | a1; slot 1: b0 | b1
  int band0, band1;
};

// X is the problem's tensor laid out with the output mode fastest: element (row, k) at
// row + Iout k, k = ia + Ia ib (the mode-0 order itself for mode 0; for modes 1 and 2 the
// copies the batch keeps in (i1, i0, i2) and (i2, i0, i1) order), so the 16 rows a load
// instruction touches are consecutive floats.
template <bool F32, int LAST_Q>
__device__ __forceinline__ void mttkrp_mfma(const float *X, int Iout, int Ia, int Ib, const double *Fa,
                                            const double *Fb, double *M, int nw, int64_t sb, int64_t se,
                                            MfmaPart &part) {
  const int t = tid(), lane = t & 63, wave = t >> 6;
  part.band0 = part.band1 = -1;
  if (wave >= nw) return;
  const int K = Ia * Ib, KS = (K + 3) / 4, KF = K / 4;
  constexpr int rs = 1;
  const int64_t S = se - sb;                 // the steps [sb, se) of the (band, k-step) space
  int64_t s0, s1;
  if constexpr (LAST_Q == 4) {
    s0 = sb + (int64_t)wave * S / nw;
    s1 = sb + (int64_t)(wave + 1) * S / nw;
  } else {
    const int64_t den = 4 * (nw - 1) + LAST_Q;
    s0 = sb + 4 * wave * S / den;
    s1 = wave + 1 >= nw ? se : sb + 4 * (wave + 1) * S / den;
  }
  const int m = lane & 15, kk = lane >> 4;
  const int inc_a = 4 % Ia, inc_b = 4 / Ia;      // k += 4 as (ia, ib) += (inc_a, inc_b), one wrap at most
  const __amdgpu_buffer_rsrc_t rx = tensor_rsrc(X, Iout * K * 4);
  const int step_bytes = __builtin_amdgcn_readfirstlane(16 * Iout);   // 4 k of one row apart
  int slot = 0;
  for (int64_t s = s0; s < s1;) {
    const int band = (int)(s / KS);
    const int64_t tb = (int64_t)band * KS, te = tb + KS, e = te < s1 ? te : s1;
    const int r0 = band * 32 + m, r1 = r0 + 16;
    const int r0c = r0 < Iout ? r0 : Iout - 1, r1c = r1 < Iout ? r1 : Iout - 1;
    const float *x0 = X + r0c * rs;
    const float *x1 = X + r1c * rs;
    int k = 4 * (int)(s - tb) + kk, ia = k % Ia, ib = k / Ia;
    f64x4 c0 = {0.0, 0.0, 0.0, 0.0}, c1 = {0.0, 0.0, 0.0, 0.0};
    auto offset = [&](int kx, int, int) { return Iout * kx; };
    auto advance = [&] {
      k += 4;
      ia += inc_a;
      ib += inc_b;
      if (ia >= Ia) {
        ia -= Ia;
        ++ib;
      }
    };
    // steps [s, ef) have all four k below K; at most one checked step follows
    const int64_t ef = tb + KF < e ? tb + KF : e;
    int n = (int)(ef - s);
    for (; n >= kSteps; n -= kSteps) {
      float xa[kSteps], xb[kSteps];
      int ka = ia, kb = ib;                            // the group's first (ia, ib), walked again below
      const int vo0 = (r0c + Iout * k) * 4, vo1 = (r1c + Iout * k) * 4;   // bytes of the group's first step
#pragma unroll
      for (int u = 0; u < kSteps; ++u) {               // all tensor loads first
        xa[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rx, vo0, step_bytes * u, 0));
        xb[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rx, vo1, step_bytes * u, 0));
        advance();
      }
      auto knext = [&] {
        ka += inc_a;
        kb += inc_b;
        if (ka >= Ia) {
          ka -= Ia;
          ++kb;
        }
      };
      if constexpr (F32) {
        f32x4 e0 = {0.f, 0.f, 0.f, 0.f}, e1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int u = 0; u < kSteps; ++u) {
          const float bv = (float)(Fa[ka * kBR + m] * Fb[kb * kBR + m]);
          e0 = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[u], bv, e0, 0, 0, 0);
          e1 = __builtin_amdgcn_mfma_f32_16x16x4f32(xb[u], bv, e1, 0, 0, 0);
          knext();
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          c0[j] = c0[j] + (double)e0[j];
          c1[j] = c1[j] + (double)e1[j];
        }
      } else {
#pragma unroll
        for (int u = 0; u < kSteps; ++u) {
          const double bv = Fa[ka * kBR + m] * Fb[kb * kBR + m];
          c0 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)xa[u], bv, c0, 0, 0, 0);
          c1 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)xb[u], bv, c1, 0, 0, 0);
          knext();
        }
      }
    }
    {
      f32x4 e0 = {0.f, 0.f, 0.f, 0.f}, e1 = {0.f, 0.f, 0.f, 0.f};   // F32: the tail's group
      for (int64_t j = ef - n; j < e; ++j) {
        const bool v = k < K;
        const int off = v ? offset(k, ia, ib) : 0;
        const double bv = v ? Fa[ia * kBR + m] * Fb[ib * kBR + m] : 0.0;
        if constexpr (F32) {
          const float av = v ? x0[off] : 0.f, aw = v ? x1[off] : 0.f;
          e0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av, (float)bv, e0, 0, 0, 0);
          e1 = __builtin_amdgcn_mfma_f32_16x16x4f32(aw, (float)bv, e1, 0, 0, 0);
        } else {
          const double av = v ? (double)x0[off] : 0.0, aw = v ? (double)x1[off] : 0.0;
          c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, c0, 0, 0, 0);
          c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(aw, bv, c1, 0, 0, 0);
        }
        advance();
      }
      if constexpr (F32) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          c0[j] = c0[j] + (double)e0[j];
          c1[j] = c1[j] + (double)e1[j];
        }
      }
    }
    if (s == tb && e == te) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = band * 32 + dtile_row<F32>(kk, j);
        if (r < Iout) M[r * kBR + m] = c0[j];
        if (r + 16 < Iout) M[(r + 16) * kBR + m] = c1[j];
      }
    } else if (slot == 0) {
      part.a0 = c0;
      part.a1 = c1;
      part.band0 = band;
      slot = 1;
    } else {
      part.b0 = c0;
      part.b1 = c1;
      part.band1 = band;
    }
    s = e;
  }
}

// the partial bands of every wave into P ([wave][slot][32 x 16]) and their ids into seg
template <bool F32>
__device__ __forceinline__ void mttkrp_mfma_partials(const MfmaPart &part, int nw, double *P, int *seg) {
  const int t = tid(), lane = t & 63, wave = t >> 6;
  if (wave >= nw) return;
  const int m = lane & 15, kk = lane >> 4;
  if (lane == 0) {
    seg[2 * wave] = part.band0;
    seg[2 * wave + 1] = part.band1;
  }
  double *p0 = P + (2 * wave) * 32 * kBR, *p1 = p0 + 32 * kBR;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int r = dtile_row<F32>(kk, j);
    if (part.band0 >= 0) {
      p0[r * kBR + m] = part.a0[j];
      p0[(r + 16) * kBR + m] = part.a1[j];
    }
    if (part.band1 >= 0) {
      p1[r * kBR + m] = part.b0[j];
      p1[(r + 16) * kBR + m] = part.b1[j];
    }
  }
}

// rows of the split bands: the partials summed in wave order
template <int BS>
__device__ __forceinline__ void mttkrp_mfma_reduce(int Iout, int nw, const double *P, const int *seg, double *M) {
  for (int e = tid(); e < Iout * kBR; e += BS) {
    const int row = e / kBR, col = e - row * kBR, band = row >> 5, rin = row & 31;
    double s = 0.0;
    bool any = false;
    for (int w = 0; w < nw; ++w)
      for (int sl = 0; sl < 2; ++sl)
        if (seg[2 * w + sl] == band) {
          const double v = P[((2 * w + sl) * 32 + rin) * kBR + col];
          s = any ? s + v : v;
          any = true;
        }
    if (any) M[e] = s;
  }
}

// The blocks running one problem (BatchDesc::V, G): this block's member index, the
// problem's exchange area and arrival counter, and the exchanges done so far (the same
// count in every member).
struct Group {
  int g, G, V;
  double *mx;
  int *ctr, *err;
  int xi;
};

// Arrival at a group exchange: every thread makes its stores visible at device scope, one
// thread adds the block's arrival and waits until `target` arrivals are counted (bounded:
// after ~2^20 polls, or once any wait has run out, it gives up and sets *err), and every
// thread's acquire fence then keeps stale L1 lines out of the reads that follow.
__device__ __forceinline__ void group_wait(int *ctr, int target, int *err) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(ctr, 1);
    for (int spin = 0; __hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target; ++spin) {
      if (spin >= (1 << 20) || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        atomicExch(err, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  __syncthreads();
  __threadfence();
}

// does the step range [sb, se) touch band b (steps [b KS, (b + 1) KS))
__device__ __forceinline__ bool piece_touches(int64_t sb, int64_t se, int KS, int b) {
  return sb < se && sb < (int64_t)(b + 1) * KS && se > (int64_t)b * KS;
}

// one mode's MTTKRP into S.M with operands S.F[0], S.F[1] on waves [0, nw).  LAST_Q = 4: the
// waves past them run `other` meanwhile; LAST_Q < 4: the last of the nw waves runs `other`
// first and then LAST_Q quarters of a share.  P holds 2 x 32 x 16 doubles per wave.  With
// grp.V > 1 the (band, k-step) space is cut in V contiguous pieces, each summed by the
// block's waves; this member computes pieces [g V / G, (g + 1) V / G), stores each piece's
// bands in the exchange area, waits for the other members and adds the pieces in order, so
// every member holds the same M, independent of G.
template <int BS, bool F32, int LAST_Q, typename Other>
__device__ __forceinline__ void mttkrp_block(const float *X, int Iout, int Ia, int Ib, BatchLds &S, double *P, int nw,
                                             Group &grp, Other &&other) {
  const int K = Ia * Ib, KS = (K + 3) / 4, NB = (Iout + 31) / 32;
  const int64_t Stot = (int64_t)NB * KS;
  const int V = grp.V, v0 = grp.g * V / grp.G, v1 = (grp.g + 1) * V / grp.G;
  double *buf = V > 1 ? grp.mx + (int64_t)(grp.xi & 1) * V * kBMax * kBR : nullptr;
  for (int v = v0; v < v1; ++v) {
    const int64_t sb = (int64_t)v * Stot / V, se = (int64_t)(v + 1) * Stot / V;
    MfmaPart part;
    if constexpr (LAST_Q == 4) {
      if ((int)(tid() >> 6) < nw) mttkrp_mfma<F32, 4>(X, Iout, Ia, Ib, S.F[0], S.F[1], S.M, nw, sb, se, part);
      else {
        part.band0 = part.band1 = -1;
        if (v == v0) other();
      }
    } else {
      if (v == v0 && (int)(tid() >> 6) == nw - 1) other();
      mttkrp_mfma<F32, LAST_Q>(X, Iout, Ia, Ib, S.F[0], S.F[1], S.M, nw, sb, se, part);
    }
    mttkrp_mfma_partials<F32>(part, nw, P, S.seg);
    __syncthreads();
    mttkrp_mfma_reduce<BS>(Iout, nw, P, S.seg, S.M);
    __syncthreads();
    if (V > 1) {
      double *slot = buf + (int64_t)v * kBMax * kBR;
      for (int e = tid(); e < Iout * kBR; e += BS)
        if (piece_touches(sb, se, KS, (e / kBR) >> 5)) slot[e] = S.M[e];
      __syncthreads();
    }
  }
  if (V > 1) {
    ++grp.xi;
    if (grp.G > 1) {
      group_wait(grp.ctr, grp.G * grp.xi, grp.err);
    } else {                                 // the block's own pieces: visible, no stale L1 lines
      __threadfence();
      __syncthreads();
      __threadfence();
    }
    for (int e = tid(); e < Iout * kBR; e += BS) {
      const int b = (e / kBR) >> 5;
      double sum = 0.0;
      bool any = false;
      for (int v = 0; v < V; ++v) {
        const int64_t sb = (int64_t)v * Stot / V, se = (int64_t)(v + 1) * Stot / V;
        if (piece_touches(sb, se, KS, b)) {
          const double x = buf[(int64_t)v * kBMax * kBR + e];
          sum = any ? sum + x : x;
          any = true;
        }
      }
      S.M[e] = sum;
    }
    __syncthreads();
  }
}

// NesterovMNLS (NesterovMNLS.hpp:181-234) for the I x R entries of one factor.  The
// gradient G = Y Zt - Wt is formed on v_mfma_f64_16x16x4f64, one 16-row tile per wave
// (tiles w, w + NW, ...), Zt = V + lambda I held as the B fragments (lane l: Zt[4q + l/16][l%16]),
// the A fragments read from the point Y in LDS.  Lane l owns the entries of its tiles at
// rows l/16 + 4j, column l % 16 (the MFMA output layout): their iterate, W~ = M + lambda A0
// and next iterate stay in registers.  Y lives in LDS in two buffers, so one barrier per
// step both publishes the next point and ANDs the stopping test (the reference's global
// stop); the step after a satisfied test is computed and discarded.  Per entry the update is
// that of the single-problem kernel (g / L through the corrected reciprocal).  Writes the
// result, after the zero-column fallback (Cpd.hpp:206-207, an exact-zero test of each
// column), to Ag (global rows, which hold the start point on entry) and to S.F[0]; returns
// the step count.
template <int BS, int RP>
__device__ __forceinline__ int nnls_block(int I, int R, BatchLds &S, double *Ag, double d1, double d2, int max_inner) {
  constexpr int NW = BS / 64, TPW = (kBMax / 16 + NW - 1) / NW, NQ = RP / 4;
  const int t = tid(), lane = t & 63, wave = t >> 6, m = lane & 15, kk = lane >> 4;
  const int T = (I + 15) / 16;
  double *Yb0 = S.F[0], *Yb1 = S.F[1];
  const double lambda = S.prm.lambda, L = S.prm.L, invL = 1.0 / L;
  double zb[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int k = 4 * q + kk;
    zb[q] = (k < R && m < R) ? S.V[k * kBR + m] + (k == m ? lambda : 0.0) : 0.0;
  }
  const bool on = m < R;
  double ac[TPW][4], an[TPW][4], wt[TPW][4];
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) {
    const int tile = wave + NW * tt;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = tile * 16 + kk + 4 * j;
      ac[tt][j] = an[tt][j] = wt[tt][j] = 0.0;
      if (tile < T && row < I) {
        const double a0 = on ? Ag[row * kBR + m] : 0.0;
        ac[tt][j] = a0;
        wt[tt][j] = on ? S.M[row * kBR + m] + lambda * a0 : 0.0;
        Yb0[row * kBR + m] = a0;                       // columns >= R: exact zeros in both
        if (!on) Yb1[row * kBR + m] = 0.0;
      }
    }
  }
  if (t < kBR) S.colnz[t] = 0;
  __syncthreads();
  // The stop test of step it (all rows, NesterovMNLS.hpp:217) is read one step later: step
  // it + 1 is computed meanwhile from the same state, and if step it passed, its result (the
  // iterate before step it, kept in ap) is returned and step it + 1 is discarded.  Each wave
  // votes with one byte; the eight bytes are read as one word.
  constexpr uint64_t kAll = NW >= 8 ? 0x0101010101010101ull : (1ull << (8 * NW)) / 255ull;
  double ap[TPW][4];
  int it = 0, b = 0;
  for (;;) {
    const double *Y = b ? Yb1 : Yb0;
    double *Yn = b ? Yb0 : Yb1;
    const double beta = it < max_inner ? S.betas[it] : 0.0;
    const uint64_t vprev = it > 0 ? S.vote[(it - 1) & 1] : 0ull;
    int pass = 1;
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt) {
      const int tile = wave + NW * tt;
      if (tile < T) {                                  // wave-uniform
        // every LDS read first and unconditional (a row past I reads its own unused LDS row;
        // its result is neither tested nor kept), then the MFMAs, then the rows
        const int ra = tile * 16 + m;                  // A fragment row (rows past I: not used)
        double af[NQ], yv[4];
#pragma unroll
        for (int q = 0; q < NQ; ++q) af[q] = Y[(ra < kBMax ? ra : 0) * kBR + 4 * q + kk];
#pragma unroll
        for (int j = 0; j < 4; ++j) yv[j] = Y[(tile * 16 + kk + 4 * j) * kBR + m];
        f64x4 g = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < NQ; ++q) g = __builtin_amdgcn_mfma_f64_16x16x4f64(af[q], zb[q], g, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int row = tile * 16 + kk + 4 * j;
          const bool valid = on && row < I;
          const double gg = g[j] - wt[tt][j];
          const double y = yv[j];
          if (valid && ((fabs(gg * y) > d1) || (gg < -d2))) pass = 0;
          const double q0 = gg * invL;
          const double tq = y - fma(fma(-q0, L, gg), invL, q0);
          const double av = tq > 0.0 ? tq : 0.0;
          if (valid) an[tt][j] = av;
          const double yn = (1.0 + beta) * av - beta * ac[tt][j];
          if (valid) Yn[row * kBR + m] = yn;
        }
      }
    }
    {
      const bool wall = __all(pass);
      if (lane == 0) ((unsigned char *)&S.vote[it & 1])[wave] = wall ? 1 : 0;
    }
    __syncthreads();
    if (it > 0 && (vprev & kAll) == kAll) {         // step it - 1 passed: its start is the result
#pragma unroll
      for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
        for (int j = 0; j < 4; ++j) ac[tt][j] = ap[tt][j];
      --it;
      break;
    }
    if (it >= max_inner) break;                     // the cap: the result is this step's start
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        ap[tt][j] = ac[tt][j];
        ac[tt][j] = an[tt][j];
      }
    b ^= 1;
    ++it;
  }
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = (wave + NW * tt) * 16 + kk + 4 * j;
      if (on && wave + NW * tt < T && row < I && ac[tt][j] != 0.0) S.colnz[m] = 1;
    }
  __syncthreads();
  int fb = 0;
  for (int k = 0; k < R; ++k) fb |= S.colnz[k] == 0;
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = (wave + NW * tt) * 16 + kk + 4 * j;
      if (on && wave + NW * tt < T && row < I) {
        double v = ac[tt][j];
        if (fb) v = 0.9 * v + 0.1 * Ag[row * kBR + m];
        Ag[row * kBR + m] = v;
        Yb0[row * kBR + m] = v;
      }
    }
  __syncthreads();
  return it;
}

// one sweep of one problem
// Member g of the problem's group: member 0 works on the problem's state and writes its
// status and log; members g > 0 on private copies of the state (identical values: every
// member takes the same steps on the same data).
template <int BS, int RP, bool F32>
__device__ __forceinline__ void sweep_problem(const BatchArgs &a, int p, int g, int sweep, BatchLds &S, double *P) {
  constexpr int NW = BS / 64;                  // the last wave solves the eigenproblem first
  const BatchDesc &d = a.desc[p];
  BatchProb &st = a.st[p];
  if (st.done) return;                                    // stopped on the threshold earlier
  const int I0 = d.I[0], I1 = d.I[1], I2 = d.I[2];
  const int R = d.R, t = threadIdx.x;
  const float *X = a.X + d.x_off;
  const bool lead = g == 0;
  const int64_t nrows_p = d.rowoff[2] + I2;
  const int64_t hrow = (d.h_off + (int64_t)(g - 1) * nrows_p) * kBR;
  double *fac = lead ? a.fac + d.f_off * kBR : a.hfac + hrow;
  double *const A0 = fac, *const A1 = fac + (int64_t)d.rowoff[1] * kBR, *const A2 = fac + (int64_t)d.rowoff[2] * kBR;
  double *gram_g = lead ? a.grams + d.g_off : a.hgrams + d.hg_off + (int64_t)(g - 1) * 3 * kBR * kBR;
  Group grp{g, d.G, d.V, a.mx + d.mx_off, a.gctr + p, a.err, st.nx};
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
    auto eig = [&] {
      const int l = t & 63;
      for (int x = l; x < kBR * kBR; x += 64) S.eig.t[x] = S.V[x];
      PTN_WAVE_SYNC();
      double amax, amin;
      eig_extremes_wave<kBR>(R, S.eig, amax, amin);
      if (l == 0) glambda_betas(amax, amin, &S.prm, a.max_inner, S.betas);
    };
    const int In = n == 0 ? I0 : n == 1 ? I1 : I2;
    // FP64 matrix cores: the eigen wave (R <= 16: microseconds) then takes 3/4 of an MTTKRP
    // share (1.76 -> 1.69 ms per 512-problem sweep); FP32 ones: the other seven waves split
    // the MTTKRP (the same share made that sweep 1.49 -> 1.53 ms)
    if constexpr (F32) mttkrp_block<BS, F32, 4>(X + (int64_t)n * a.xcopy, In, Ia, Ib, S, P, NW - 1, grp, eig);
    else mttkrp_block<BS, F32, 3>(X + (int64_t)n * a.xcopy, In, Ia, Ib, S, P, NW, grp, eig);
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
  double *oldf = lead ? a.old + d.f_off * kBR : a.hold + hrow, *accf = lead ? a.acc + d.f_off * kBR : a.hacc + hrow;
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
      mttkrp_block<BS, F32, 4>(X + 2 * a.xcopy, I2, I0, I1, S, P, NW, grp, [] {});
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
  if (t == 0 && lead) {
    st.nx = grp.xi;
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

// Partial MTTKRP bands (2 x 32 x 16 doubles per wave): in LDS for 512-thread blocks (one
// block per CU), in the block's slice of global scratch for 256-thread blocks, whose LDS then
// fits two blocks per CU (two waves per SIMD; the launch bound keeps the registers within it).
template <int BS, bool F32>
__global__ __launch_bounds__(BS, 2) void k_batch_sweep(BatchArgs a) {
  __shared__ BatchLds S;
  __shared__ double Plds[BS == 512 ? (BS / 64) * 2 * 32 * kBR : 1];
  double *P = BS == 512 ? Plds : a.pscratch + (int64_t)blockIdx.x * kBatchPartDoubles256;
  __shared__ int sp;
  const int sweep = *a.sweep_ctr;
  // The queue holds the members of each problem's group next to each other (largest work
  // first): at most one group is incomplete at any time, so a member waiting for the others
  // waits only for blocks that are free or about to be (the group size never exceeds the grid).
  for (;;) {
    if (threadIdx.x == 0) sp = atomicAdd(&a.queue[sweep], 1);
    __syncthreads();
    const int q = sp;
    __syncthreads();
    if (q >= a.nitems) break;
    const int item = a.order[q], p = item >> 3, g = item & 7;   // (problem, member)
    switch ((a.desc[p].R + 3) >> 2) {
      case 1: sweep_problem<BS, 4, F32>(a, p, g, sweep, S, P); break;
      case 2: sweep_problem<BS, 8, F32>(a, p, g, sweep, S, P); break;
      case 3: sweep_problem<BS, 12, F32>(a, p, g, sweep, S, P); break;
      default: sweep_problem<BS, 16, F32>(a, p, g, sweep, S, P); break;
    }
  }
}

// The copies of one problem's tensor with mode 1 / mode 2 fastest, written in their own
// order: X1[i1 + I1 (i0 + I0 i2)] and X2[i2 + I2 (i0 + I0 i1)] from X[i0 + I0 (i1 + I1 i2)].
__global__ void k_batch_permute(const float *__restrict__ X, int I0, int I1, int I2, float *__restrict__ X1,
                                float *__restrict__ X2) {
  const int64_t n = (int64_t)I0 * I1 * I2;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    if (blockIdx.y == 0) {
      const int64_t i1 = e % I1, r = e / I1, i0 = r % I0, i2 = r / I0;
      X1[e] = X[i0 + I0 * (i1 + I1 * i2)];
    } else {
      const int64_t i2 = e % I2, r = e / I2, i0 = r % I0, i1 = r / I0;
      X2[e] = X[i0 + I0 * (i1 + I1 * i2)];
    }
  }
}

// the problem's state (factor rows, previous rows, Grams) into every helper's copy
__global__ void k_batch_sync_helpers(BatchArgs a) {
  const BatchDesc &d = a.desc[blockIdx.x];
  const int64_t nr = (int64_t)(d.I[0] + d.I[1] + d.I[2]) * kBR, ng = 3 * kBR * kBR;
  for (int g = 1; g < d.G; ++g) {
    const int64_t h = (d.h_off + (int64_t)(g - 1) * (d.I[0] + d.I[1] + d.I[2])) * kBR;
    for (int64_t e = threadIdx.x; e < nr; e += blockDim.x) {
      a.hfac[h + e] = a.fac[d.f_off * kBR + e];
      a.hold[h + e] = a.old[d.f_off * kBR + e];
    }
    for (int64_t e = threadIdx.x; e < ng; e += blockDim.x) a.hgrams[d.hg_off + (g - 1) * ng + e] = a.grams[d.g_off + e];
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
  if (a.arith_f32) {
    if (block == 512) k_batch_sweep<512, true><<<grid, 512, 0, s>>>(a);
    else k_batch_sweep<256, true><<<grid, 256, 0, s>>>(a);
  } else {
    if (block == 512) k_batch_sweep<512, false><<<grid, 512, 0, s>>>(a);
    else k_batch_sweep<256, false><<<grid, 256, 0, s>>>(a);
  }
}

void batch_next_sweep(hipStream_t s, int *ctr) { k_batch_next<<<1, 64, 0, s>>>(ctr); }

void batch_sync_helpers(hipStream_t s, const BatchArgs &a) { k_batch_sync_helpers<<<a.nprob, 256, 0, s>>>(a); }

void batch_permute(hipStream_t s, const float *X, int I0, int I1, int I2, float *X1, float *X2) {
  const int64_t n = (int64_t)I0 * I1 * I2;
  int64_t g = (n + 255) / 256;
  if (g > 2048) g = 2048;
  k_batch_permute<<<dim3((unsigned)g, 2), 256, 0, s>>>(X, I0, I1, I2, X1, X2);
}

}  // namespace ptn
