// FP32-stream tensor contractions on gfx950 matrix cores.
//
// The tensor is read in FP32 (half the bytes of the FP64 path) and multiplied on
// the matrix cores either as bf16x3 split products (x = x_hi + x_lo, w = w_hi +
// w_lo in bf16; x_hi w_hi + x_hi w_lo + x_lo w_hi accumulated in FP32 by
// v_mfma_f32_32x32x16_bf16) or in plain FP32 (v_mfma_f32_32x32x2_f32).  The
// Khatri-Rao operand W is formed once per contraction from the FP64 factor rows
// and stored in MFMA fragment order, so a lane fetches its 8 bf16 with one
// 16-byte load.
//
// Shapes (same as the FP64 kernels):
//   gemm_cm_f32:  out[m][r] = sum_k X[m + M*k] W(k)[r]   (contracted index strided)
//   gemm_ck_f32:  T[c][r]   = sum_a X[a + L*c] W(a)[r]   (contracted index contiguous)
//
// MFMA 32x32 layouts (checked on gfx950 with a probe): lane l, kg = l/32, i = l%32
//   A (32 x K): A[i][8kg + j] (bf16: j < 8; f32 32x32x2: K = 2, A[i][kg])
//   B (K x 32): B[8kg + j][i]
//   D (32 x 32): D[8(v/4) + 4kg + v%4][i], v = 0..15
// For FP32 MFMA step q (q < 8) the two k of the instruction are the lane's
// (8kg + q): both operands use the same permutation of k, so the sum is unchanged.
#include "ptn_internal.hpp"

namespace ptn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x8 __attribute__((ext_vector_type(8)));

namespace {

struct KrArgs32 {
  int nm;
  uint32_t dim[kMaxOrder];
  const double *fac[kMaxOrder];
};

__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

// W fragments, one bf16x8 per (16-wide k step s16, column tile t, lane l):
//   c = 32 t + l%32 and element j holds W(k)[c] (0 for k >= K) with
//   kperm 0:  k = 16 s16 + 8 (l/32) + j                       (lane pairs split 8 + 8)
//   kperm 1:  k = 32 (s16/2) + 16 (l/32) + 8 (s16%2) + j       (lane pairs split 16 + 16)
// W(k)[c] = prod_q fac_q[d_q(k)][c]; wf: the same fragments in FP32 (one f32x8 per
// (s16, t, lane), element j the FP32 value of the bf16 pair's W(k)), so that the FP32
// contractions fetch a step's operand with one 32-byte load per column tile.
// Column tiles t >= tsplit hold a second operand w2 (column c - 32 tsplit of its factors):
// one contraction then serves two sets of factors.
__global__ void k_pack_w(KrArgs32 w1, KrArgs32 w2, int tsplit, int64_t K, int64_t ksteps, int NT, int ldr,
                         int kperm, bf16x8 *whi, bf16x8 *wlo, float *wf) {
  const int64_t n = ksteps * NT * 64;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int l = (int)(e % 64);
    const int64_t st = e / 64;
    const int t = (int)(st % NT);
    const int64_t s16 = st / NT;
    const int c = 32 * t + (l % 32);
    const bool second = t >= tsplit;
    const KrArgs32 &w = second ? w2 : w1;
    const int cf = second ? c - 32 * tsplit : c;
    const int64_t kbase = kperm ? 32 * (s16 / 2) + 16 * (l / 32) + 8 * (s16 % 2) : 16 * s16 + 8 * (l / 32);
    bf16x8 h, lo;
    f32x8 fv;
    // the eight k of the fragment mode by mode, so that each mode's eight row loads are
    // independent and in flight together (same left-to-right products; an entry past K or
    // past the rank reads row 0 / column 0 and is zeroed)
    const bool cv = cf < ldr;
    const int cc = cv ? cf : 0;
    uint32_t rem[8];
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      rem[j] = kbase + j < K ? (uint32_t)(kbase + j) : 0u;
      v[j] = 1.0;
    }
    for (int q = 0; q < w.nm; ++q) {
      const double *fq = w.fac[q];
      const uint32_t dq = w.dim[q];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint32_t d = rem[j] % dq;
        rem[j] /= dq;
        v[j] *= fq[(int64_t)d * ldr + cc];
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      float f = (kbase + j < K && cv) ? (float)v[j] : 0.f;
      __bf16 a, b;
      split_bf16(f, a, b);
      h[j] = a;
      lo[j] = b;
      fv[j] = f;
    }
    if (whi) { whi[e] = h; wlo[e] = lo; }
    if (wf) ((f32x8 *)wf)[e] = fv;
  }
}

template <int MODE>
__device__ __forceinline__ void mma_step(const float (&x)[8], const bf16x8 &bhi, const bf16x8 &blo, const f32x8 &bf,
                                         f32x16 &acc) {
  if constexpr (MODE == kGemmBf16x3) {
    bf16x8 ahi, alo;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      __bf16 h, l;
      split_bf16(x[j], h, l);
      ahi[j] = h;
      alo[j] = l;
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bhi, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, blo, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi, acc, 0, 0, 0);
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(x[q], bf[q], acc, 0, 0, 0);
  }
}

// ---------------------------------------------------------------------------
// gemm_cm_f32: each wave owns 64 consecutive rows (2 tiles of 32) and the block's
// k-range.  Each 16-wide k step loads its operand fragments and the 2 x 8 tensor
// values of every lane (unconditional loads from clamped addresses), converts and
// multiplies; occupancy hides the load latency, except for FP32 products at up to 64
// columns, which issue the next step's loads before the current step's MFMAs (two
// register sets).  The 64 x 32 result of each column tile
// leaves through LDS as 16-byte non-temporal row stores: scattered 4-byte stores of the
// MFMA layout cost several times their bytes next to the tensor stream.  CAP (bf16x3 at
// 32 or 64 columns, FP32 products at 64): the staging area is padded so that at most two
// blocks share a CU; the strided tensor stream runs faster with 8 waves per CU than with
// the 12-20 the registers allow (800^3 left node, 32 columns: 3-10%; 200^4 order-4 node,
// 64 columns split 13 ways: 1.113 ms against 1.234 ms bf16x3, 2.15 against 2.22 ms FP32).
// ---------------------------------------------------------------------------
// A raw buffer over `bytes` bytes from p (p uniform; its halves taken to scalar registers)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const float *p, uint32_t bytes) {
  const uint64_t a = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  return __builtin_amdgcn_make_buffer_rsrc((void *)(((uint64_t)hi << 32) | lo), (short)0,
                                           __builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}
__device__ __forceinline__ float load_nt(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 2));   // 2: nt
}

template <int MODE, int NT, bool CAP, bool BUF, bool KEEP>
__global__ __launch_bounds__(256, NT >= 4 || MODE == kGemmFp32 ? 2 : 1) void k_gemm_cm_f32(const float *__restrict__ X, int64_t M, int64_t K,
                                                     int64_t kchunk, const bf16x8 *__restrict__ whi,
                                                     const bf16x8 *__restrict__ wlo, const float *__restrict__ wf,
                                                     float *__restrict__ out) {
  constexpr int RPF = 32 * NT;
  constexpr int PAD = CAP ? 2048 : 0;                    // floats per wave: 66.5 KB per block
  __shared__ float stg[4][64 * 33 + PAD];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = lane & 31, kg = lane >> 5;
  const int64_t mb = ((int64_t)blockIdx.x * 4 + wave) * 64;
  if (mb >= M) return;                                   // no block-wide barrier below
  const int64_t kbeg = (int64_t)blockIdx.y * kchunk;
  const int64_t kend = kbeg + kchunk < K ? kbeg + kchunk : K;

  f32x16 acc[2][NT];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[s][t][v] = 0.f;

  // A missing row only fills its own (discarded) row of D, and k >= K meets the zero
  // rows of the packed operand.  Splits are multiples of 16 columns, so a step never
  // crosses into the next split.
  const int64_t row0 = mb + i < M ? mb + i : M - 1, row1 = mb + 32 + i < M ? mb + 32 + i : M - 1;
  const float *p0 = X + row0;
  const float *p1 = X + row1;
  auto step = [&](int64_t k0, const float (&xc)[2][8]) {
    const int64_t ks = k0 >> 4;
    bf16x8 bh[NT], bl[NT];
    f32x8 bf[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if constexpr (MODE == kGemmBf16x3) {
        bh[t] = whi[(ks * NT + t) * 64 + lane];
        bl[t] = wlo[(ks * NT + t) * 64 + lane];
      } else {
        bf[t] = ((const f32x8 *)wf)[(ks * NT + t) * 64 + lane];
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      mma_step<MODE>(xc[0], bh[t], bl[t], bf[t], acc[0][t]);
      mma_step<MODE>(xc[1], bh[t], bl[t], bf[t], acc[1][t]);
    }
  };
  int64_t k0 = kbeg;
  if constexpr (BUF) {
    // whole 16-k steps through a raw buffer over the block's k range (its byte size fits 32
    // bits, checked by the launcher): the lane's row and k group in one offset register, the
    // step's row in a scalar one, no 64-bit address per load
    const __amdgpu_buffer_rsrc_t rx = raw_rsrc(X + kbeg * M, (uint32_t)((kend - kbeg) * M * 4));
    const uint32_t vo0 = (uint32_t)((row0 + 8 * kg * M) * 4), vo1 = (uint32_t)((row1 + 8 * kg * M) * 4);
    const uint32_t rowb = __builtin_amdgcn_readfirstlane((uint32_t)(M * 4));
    if constexpr (MODE == kGemmFp32 && NT <= 2) {
      // FP32 products: 16 MFMAs per column tile and step leave the matrix cores the bound, so
      // the next step's tensor values and operand fragments are loaded into a second set of
      // registers while this step multiplies (same products, same order; 200^4 order-4 node
      // at 64 columns: 2.42 -> 2.22 ms)
      struct Buf {
        float x[2][8];
        f32x8 b[NT];
      };
      auto load = [&](int64_t k, Buf &u) {
        const uint32_t sb = __builtin_amdgcn_readfirstlane((uint32_t)((k - kbeg) * M * 4));
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          u.x[0][j] = load_nt(rx, vo0, sb + j * rowb);
          u.x[1][j] = load_nt(rx, vo1, sb + j * rowb);
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) u.b[t] = ((const f32x8 *)wf)[((k >> 4) * NT + t) * 64 + lane];
      };
      auto mma = [&](const Buf &u) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          mma_step<MODE>(u.x[0], bf16x8{}, bf16x8{}, u.b[t], acc[0][t]);
          mma_step<MODE>(u.x[1], bf16x8{}, bf16x8{}, u.b[t], acc[1][t]);
        }
      };
      int n = __builtin_amdgcn_readfirstlane((int)((kend - k0) / 16));   // whole steps of the range
      Buf b0, b1;
      if (n & 1) {                                       // an odd first step alone
        load(k0, b0);
        mma(b0);
        k0 += 16;
        --n;
      }
      if (n > 0) {
        // pairs of steps; the loads of the step after the last pair re-read the last step
        // (unconditional, so the loop has no branch between a load and its use)
        load(k0, b0);
        for (int j = 0; j < n; j += 2) {
          load(k0 + 16 * (j + 1), b1);
          __builtin_amdgcn_sched_barrier(0);
          mma(b0);
          load(k0 + 16 * (j + 2 < n ? j + 2 : n - 1), b0);
          __builtin_amdgcn_sched_barrier(0);
          mma(b1);
        }
        k0 += 16 * n;
      }
    } else {
      for (; k0 + 16 <= kend; k0 += 16) {
        const uint32_t sb = __builtin_amdgcn_readfirstlane((uint32_t)((k0 - kbeg) * M * 4));
        float xc[2][8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          xc[0][j] = load_nt(rx, vo0, sb + j * rowb);
          xc[1][j] = load_nt(rx, vo1, sb + j * rowb);
        }
        step(k0, xc);
      }
    }
  }
  for (; k0 < kend; k0 += 16) {
    float xc[2][8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int64_t k = k0 + 8 * kg + j;
      const int64_t kc = k < K ? k : K - 1;
      xc[0][j] = __builtin_nontemporal_load(p0 + M * kc);
      xc[1][j] = __builtin_nontemporal_load(p1 + M * kc);
    }
    step(k0, xc);
  }
  float *o = out + (int64_t)blockIdx.y * M * RPF;
  float *st = stg[wave];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      st[(8 * (v / 4) + 4 * kg + (v % 4)) * 33 + i] = acc[0][t][v];
      st[(32 + 8 * (v / 4) + 4 * kg + (v % 4)) * 33 + i] = acc[1][t][v];
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = 64 * q + lane;                       // 16-byte chunk of the 64 x 32 tile
      const int r = e >> 3, c4 = 4 * (e & 7);
      const int64_t row = mb + r;
      const f32x4 val = {st[r * 33 + c4], st[r * 33 + c4 + 1], st[r * 33 + c4 + 2], st[r * 33 + c4 + 3]};
      if (row < M) {
        // KEEP (split partial sums): cached stores, read back soon by the split reduction;
        // a whole node goes out non-temporally
        f32x4 *dst = (f32x4 *)(o + row * RPF + 32 * t + c4);
        if constexpr (KEEP) *dst = val;
        else __builtin_nontemporal_store(val, dst);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// ---------------------------------------------------------------------------
// gemm_ck_f32: each wave owns 64 fibers (2 MFMA row tiles of 32) and walks the
// contraction in 32-element steps: lane (i, kg) reads 16 consecutive elements (64
// bytes) of its fiber, so each step consumes whole 128-byte lines.  Operands use the
// kperm = 1 k order of k_pack_w.  FUSE: the fibers of a wave share one index of the
// last mode (c = i1 + I1 * i2, 64 consecutive i1 per wave) and the epilogue multiplies
// each fiber row by the FP64 factor row A1[i1] and sums them: the right node of the
// order-3 dimension tree (X x_0 A_0, then the TTV with A_1, DimTrees.hpp:668-800) in
// one pass, written as one partial row per (i2, i1-chunk).
// ---------------------------------------------------------------------------
template <int MODE, int NT, bool VEC, bool FUSE>
__global__ __launch_bounds__(256) void k_gemm_ck_f32(const float *__restrict__ X, int64_t L, int64_t C, int64_t I1,
                                                     const bf16x8 *__restrict__ whi, const bf16x8 *__restrict__ wlo,
                                                     const float *__restrict__ wf, float *__restrict__ T,
                                                     const double *__restrict__ A1, int ldr,
                                                     double *__restrict__ part) {
  constexpr int RPF = 32 * NT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = lane & 31, kg = lane >> 5;
  const int64_t wid = (int64_t)blockIdx.x * 4 + wave;
  int64_t cb, nfib, i1b = 0, i2 = 0;
  if (FUSE) {
    const int64_t wpf = (I1 + 63) / 64;
    i2 = wid / wpf;
    i1b = (wid % wpf) * 64;
    cb = i1b + I1 * i2;
    nfib = I1 - i1b < 64 ? I1 - i1b : 64;
    if (i2 * I1 >= C) return;
  } else {
    cb = wid * 64;
    if (cb >= C) return;
    nfib = C - cb < 64 ? C - cb : 64;
  }

  f32x16 acc[2][NT];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[s][t][v] = 0.f;

  // Unconditional loads and nothing zeroed: a missing fiber reads fiber cb and only
  // fills its own (discarded) row of D; elements a >= L meet the zero rows of the
  // packed operand.  VEC (L % 4 == 0) reads whole 32-element steps: past the end of a
  // fiber that is the next fiber, and past the last fiber the allocation's zero pad
  // (at least 32 elements).  The next step's loads are clamped to the last step, so
  // every iteration takes the same path.
  const float *f0 = X + L * (cb + (i < nfib ? i : 0));
  const float *f1 = X + L * (cb + (32 + i < nfib ? 32 + i : 0));
  const int64_t alast = ((L - 1) / 32) * 32;
  float xn[2][16];
  auto load = [&](int64_t a0, float (&dst)[2][16]) {
    const int64_t a = (a0 < alast ? a0 : alast) + 16 * kg;
    if (VEC) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const f32x4 u = __builtin_nontemporal_load((const f32x4 *)(f0 + a + 4 * q));
        const f32x4 w = __builtin_nontemporal_load((const f32x4 *)(f1 + a + 4 * q));
#pragma unroll
        for (int j = 0; j < 4; ++j) { dst[0][4 * q + j] = u[j]; dst[1][4 * q + j] = w[j]; }
      }
    } else {
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int64_t ac = a + j < L ? a + j : L - 1;
        dst[0][j] = f0[ac];
        dst[1][j] = f1[ac];
      }
    }
  };
  load(0, xn);
  for (int64_t a0 = 0; a0 < L; a0 += 32) {
    float xc[2][16];
#pragma unroll
    for (int j = 0; j < 16; ++j) { xc[0][j] = xn[0][j]; xc[1][j] = xn[1][j]; }
    const int64_t s32 = a0 >> 5;
    // this step's packed operand first: memory counters retire in order, so loads
    // issued after the next step's tensor loads would make the MFMAs wait for them
    bf16x8 bh[2][NT], bl[2][NT];
    f32x8 bf[2][NT];
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        if constexpr (MODE == kGemmBf16x3) {
          bh[s2][t] = whi[((s32 * 2 + s2) * NT + t) * 64 + lane];
          bl[s2][t] = wlo[((s32 * 2 + s2) * NT + t) * 64 + lane];
        } else {
          bf[s2][t] = ((const f32x8 *)wf)[((s32 * 2 + s2) * NT + t) * 64 + lane];
        }
      }
    load(a0 + 32, xn);
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2) {
      float x0[8], x1[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) { x0[j] = xc[0][8 * s2 + j]; x1[j] = xc[1][8 * s2 + j]; }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        // FP32 path: lane (c, kg) supplies W[a0 + 16 kg + 8 s2 + q][c] at MFMA step q
        mma_step<MODE>(x0, bh[s2][t], bl[s2][t], bf[s2][t], acc[0][t]);
        mma_step<MODE>(x1, bh[s2][t], bl[s2][t], bf[s2][t], acc[1][t]);
      }
    }
  }
  if (!FUSE) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t f = 32 * s + 8 * (v / 4) + 4 * kg + (v % 4);
        if (f < nfib) {
#pragma unroll
          for (int t = 0; t < NT; ++t) T[(cb + f) * RPF + 32 * t + i] = acc[s][t][v];
        }
      }
    return;
  }
  // fused TTV with A1 over the wave's fibers (FP64 accumulation)
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int col = 32 * t + i;
    double sum = 0.0;
    if (col < ldr) {
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int64_t f = 32 * s + 8 * (v / 4) + 4 * kg + (v % 4);
          if (f < nfib) sum = fma((double)acc[s][t][v], A1[(i1b + f) * ldr + col], sum);
        }
    }
    sum += __shfl_xor(sum, 32, 64);
    if (kg == 0 && col < ldr) part[wid * ldr + col] = sum;
  }
}

// ---------------------------------------------------------------------------
// gemm_ck_lds: the same contraction as gemm_ck_f32 (L % 4 == 0) with the tensor staged
// through LDS.  A block owns 256 consecutive fibers; per 32-element step each group of
// 8 consecutive threads reads one whole 128-byte line of one fiber (global requests are
// full lines), the 256 x 32 tile is written to LDS (rows padded to 36 floats, which
// keeps the 16-byte reads and writes free of bank conflicts), and each wave reads its
// 64 fibers back in the MFMA layout.  The next step's lines are prefetched into
// registers while the current step is multiplied.  FUSE as in gemm_ck_f32, with the
// block's fibers inside one index of the last mode.
// ---------------------------------------------------------------------------
template <int MODE, int NT, bool FUSE>
__global__ __launch_bounds__(256) void k_gemm_ck_lds(const float *__restrict__ X, int64_t L, int64_t C, int64_t I1,
                                                     const bf16x8 *__restrict__ whi, const bf16x8 *__restrict__ wlo,
                                                     const float *__restrict__ wf, float *__restrict__ T,
                                                     const double *__restrict__ A1, int ldr,
                                                     double *__restrict__ part) {
  constexpr int RPF = 32 * NT, TF = 256, LDP = 36;
  // unfused bf16x3: LDS padded to at most two blocks per CU (55 KB each), which measured
  // faster beside the solver kernels of the other streams; the fused epilogue is not
  constexpr int PAD = (MODE == kGemmBf16x3 && !FUSE && NT == 1) ? 4608 : 0;
  __shared__ float tile[TF * LDP + PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i = lane & 31, kg = lane >> 5;
  int64_t cb, nfb, i1b = 0, i2 = 0;
  if (FUSE) {
    const int64_t bpf = (I1 + TF - 1) / TF;
    i2 = blockIdx.x / bpf;
    i1b = (blockIdx.x % bpf) * TF;
    if (i2 * I1 >= C) return;
    cb = i1b + I1 * i2;
    nfb = I1 - i1b < TF ? I1 - i1b : TF;
  } else {
    cb = (int64_t)blockIdx.x * TF;
    nfb = C - cb < TF ? C - cb : TF;
  }
  // global side: instruction q reads fiber 32 q + tid / 8, elements 4 (tid % 8) .. + 3
  const int k4 = 4 * (tid & 7);
  const float *src[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int fl = 32 * q + (tid >> 3);
    src[q] = X + L * (cb + (fl < nfb ? fl : 0)) + k4;
  }
  const int64_t alast = ((L - 1) / 32) * 32;
  f32x4 pre[8];
  auto fetch = [&](int64_t a0) {
    const int64_t a = a0 < alast ? a0 : alast;
#pragma unroll
    for (int q = 0; q < 8; ++q) pre[q] = __builtin_nontemporal_load((const f32x4 *)(src[q] + a));
  };

  f32x16 acc[2][NT];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[s][t][v] = 0.f;

  fetch(0);
  for (int64_t a0 = 0; a0 < L; a0 += 32) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) *(f32x4 *)(tile + (32 * q + (tid >> 3)) * LDP + k4) = pre[q];
    __syncthreads();
    const int64_t s32 = a0 >> 5;
    bf16x8 bh[2][NT], bl[2][NT];
    f32x8 bf[2][NT];
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        if constexpr (MODE == kGemmBf16x3) {
          bh[s2][t] = whi[((s32 * 2 + s2) * NT + t) * 64 + lane];
          bl[s2][t] = wlo[((s32 * 2 + s2) * NT + t) * 64 + lane];
        } else {
          bf[s2][t] = ((const f32x8 *)wf)[((s32 * 2 + s2) * NT + t) * 64 + lane];
        }
      }
    fetch(a0 + 32);
    // keeps the next step's loads here, ahead of this step's MFMAs: without it the compiler
    // issues them at the top of the next iteration, right before the LDS write that waits
    // for them (1000^3 R = 64 sweep: 346 -> 358 per second)
    __builtin_amdgcn_sched_barrier(0);
    float xc[2][16];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const float *row = tile + (64 * wave + 32 * s + i) * LDP + 16 * kg;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const f32x4 u = *(const f32x4 *)(row + 4 * q);
#pragma unroll
        for (int j = 0; j < 4; ++j) xc[s][4 * q + j] = u[j];
      }
    }
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2) {
      float x0[8], x1[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) { x0[j] = xc[0][8 * s2 + j]; x1[j] = xc[1][8 * s2 + j]; }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        mma_step<MODE>(x0, bh[s2][t], bl[s2][t], bf[s2][t], acc[0][t]);
        mma_step<MODE>(x1, bh[s2][t], bl[s2][t], bf[s2][t], acc[1][t]);
      }
    }
  }
  const int64_t wb = 64 * wave;               // the wave's first fiber within the block
  if (!FUSE) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t f = wb + 32 * s + 8 * (v / 4) + 4 * kg + (v % 4);
        if (f < nfb) {
#pragma unroll
          for (int t = 0; t < NT; ++t) T[(cb + f) * RPF + 32 * t + i] = acc[s][t][v];
        }
      }
    return;
  }
  if (wb >= nfb) return;                      // wave-uniform
  const int64_t wpf = (I1 + 63) / 64;
  const int64_t chunk = i2 * wpf + (i1b + wb) / 64;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int col = 32 * t + i;
    double sum = 0.0;
    if (col < ldr) {
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int64_t f = wb + 32 * s + 8 * (v / 4) + 4 * kg + (v % 4);
          if (f < nfb) sum = fma((double)acc[s][t][v], A1[(i1b + f) * ldr + col], sum);
        }
    }
    sum += __shfl_xor(sum, 32, 64);
    if (kg == 0 && col < ldr) part[chunk * ldr + col] = sum;
  }
}

// M[i2][r] = sum over the i1-chunks of the fused right node, in chunk order.
__global__ void k_reduce_chunks(const double *__restrict__ part, int64_t I2, int64_t wpf, int ldr,
                                double *__restrict__ M) {
  const int64_t n = I2 * ldr;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i2 = e / ldr;
    const int r = (int)(e % ldr);
    double s = 0.0;
    for (int64_t c = 0; c < wpf; ++c) s += part[(i2 * wpf + c) * ldr + r];
    M[e] = s;
  }
}

__global__ void k_reduce_splits_f32(const float *__restrict__ partials, int splits, int64_t n, float *__restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int q = 0; q < splits; ++q) s += partials[q * n + e];
    out[e] = s;
  }
}
// the same sums four at a time (n % 4 == 0): 16-byte loads of every split, in split order
__global__ void k_reduce_splits_f32x4(const f32x4 *__restrict__ partials, int splits, int64_t n4,
                                      f32x4 *__restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n4; e += (int64_t)gridDim.x * blockDim.x) {
    f32x4 s = {0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < splits; ++q) {
      const f32x4 v = __builtin_nontemporal_load(partials + q * n4 + e);
      s[0] += v[0];
      s[1] += v[1];
      s[2] += v[2];
      s[3] += v[3];
    }
    out[e] = s;
  }
}

KrArgs32 to_args32(const KrSpec &w) {
  KrArgs32 a{};
  a.nm = w.nm;
  for (int j = 0; j < w.nm; ++j) { a.dim[j] = (uint32_t)w.dim[j]; a.fac[j] = w.fac[j]; }
  return a;
}

}  // namespace

int64_t wpack_elems(int64_t K, int RPF) { return ((K + 31) / 32) * 2 * (RPF / 32) * 64; }

void pack_w(hipStream_t s, const KrSpec &w, const KrSpec *w2, int64_t K, int RPF, int ldr, int kperm, void *whi,
            void *wlo, float *wf) {
  const int NT = RPF / 32;
  const int tsplit = w2 ? NT / 2 : NT;
  const int64_t ksteps = ((K + 31) / 32) * 2;
  const int64_t n = ksteps * NT * 64;
  int64_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  k_pack_w<<<(unsigned)g, 256, 0, s>>>(to_args32(w), to_args32(w2 ? *w2 : w), tsplit, K, ksteps, NT, ldr, kperm,
                                       (bf16x8 *)whi, (bf16x8 *)wlo, wf);
}

void gemm_cm_f32(hipStream_t s, GemmMode mode, const float *x32, int64_t M, int64_t K, int RPF, int splits,
                 const void *whi, const void *wlo, const float *wf, float *out) {
  int64_t kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + 15) / 16 * 16;
  dim3 grid((unsigned)((M + 255) / 256), (unsigned)splits);
  const bf16x8 *h = (const bf16x8 *)whi, *l = (const bf16x8 *)wlo;
  // raw-buffer loads when a block's k range of the tensor spans less than 4 GB
  const bool buf = (kchunk + 16) * M * 4 < ((int64_t)1 << 32);
#define PTN_CM32K(MODEV, NTV, CAPV, KEEPV)                                                                  \
  do {                                                                                                      \
    if (buf) k_gemm_cm_f32<MODEV, NTV, CAPV, true, KEEPV><<<grid, 256, 0, s>>>(x32, M, K, kchunk, h, l, wf, out); \
    else k_gemm_cm_f32<MODEV, NTV, CAPV, false, KEEPV><<<grid, 256, 0, s>>>(x32, M, K, kchunk, h, l, wf, out); \
  } while (0)
#define PTN_CM32(MODEV, NTV, CAPV)              \
  do {                                          \
    if (splits > 1) PTN_CM32K(MODEV, NTV, CAPV, true); \
    else PTN_CM32K(MODEV, NTV, CAPV, false);   \
  } while (0)
  if (mode == kGemmBf16x3) {
    if (RPF == 32) PTN_CM32(kGemmBf16x3, 1, true);
    else if (RPF == 64) PTN_CM32(kGemmBf16x3, 2, true);
    else PTN_CM32(kGemmBf16x3, 4, false);
  } else {
    if (RPF == 32) PTN_CM32(kGemmFp32, 1, false);
    else if (RPF == 64) PTN_CM32(kGemmFp32, 2, true);
    else PTN_CM32(kGemmFp32, 4, false);
  }
#undef PTN_CM32
#undef PTN_CM32K
}

void gemm_ck_f32(hipStream_t s, GemmMode mode, const float *x32, int64_t L, int64_t C, int RPF, const void *whi,
                 const void *wlo, const float *wf, float *T, int64_t fuse_I1, const double *A1, int ldr,
                 double *part) {
  const bool fuse = fuse_I1 > 0;
  const bf16x8 *h = (const bf16x8 *)whi, *l = (const bf16x8 *)wlo;
  const bool vec = (L % 4) == 0;
  if (vec) {
    const int64_t blocks = fuse ? (C / fuse_I1) * ((fuse_I1 + 255) / 256) : (C + 255) / 256;
    dim3 g2((unsigned)blocks);
#define PTN_CKL(MODEV, NTV, FUSEV) \
  k_gemm_ck_lds<MODEV, NTV, FUSEV><<<g2, 256, 0, s>>>(x32, L, C, fuse_I1, h, l, wf, T, A1, ldr, part)
    if (mode == kGemmBf16x3) {
      if (RPF == 32) { if (fuse) PTN_CKL(kGemmBf16x3, 1, true); else PTN_CKL(kGemmBf16x3, 1, false); }
      else           { if (fuse) PTN_CKL(kGemmBf16x3, 2, true); else PTN_CKL(kGemmBf16x3, 2, false); }
    } else {
      if (RPF == 32) { if (fuse) PTN_CKL(kGemmFp32, 1, true); else PTN_CKL(kGemmFp32, 1, false); }
      else           { if (fuse) PTN_CKL(kGemmFp32, 2, true); else PTN_CKL(kGemmFp32, 2, false); }
    }
#undef PTN_CKL
    return;
  }
  const int64_t waves = fuse ? (C / fuse_I1) * ((fuse_I1 + 63) / 64) : (C + 63) / 64;
  dim3 grid((unsigned)((waves + 3) / 4));
#define PTN_CK32(MODEV, NTV, VECV, FUSEV)                                                                  \
  k_gemm_ck_f32<MODEV, NTV, VECV, FUSEV><<<grid, 256, 0, s>>>(x32, L, C, fuse_I1, h, l, wf, T, A1, ldr, part)
#define PTN_CK32_F(MODEV, NTV, VECV) \
  { if (fuse) PTN_CK32(MODEV, NTV, VECV, true); else PTN_CK32(MODEV, NTV, VECV, false); }
#define PTN_CK32_V(MODEV, NTV) \
  { if (vec) PTN_CK32_F(MODEV, NTV, true) else PTN_CK32_F(MODEV, NTV, false) }
  if (mode == kGemmBf16x3) {
    if (RPF == 32) PTN_CK32_V(kGemmBf16x3, 1) else PTN_CK32_V(kGemmBf16x3, 2)
  } else {
    if (RPF == 32) PTN_CK32_V(kGemmFp32, 1) else PTN_CK32_V(kGemmFp32, 2)
  }
#undef PTN_CK32_V
#undef PTN_CK32_F
#undef PTN_CK32
}

void reduce_chunks(hipStream_t s, const double *part, int64_t I2, int64_t wpf, int ldr, double *M) {
  int64_t g = (I2 * ldr + 255) / 256;
  if (g > 4096) g = 4096;
  k_reduce_chunks<<<(unsigned)g, 256, 0, s>>>(part, I2, wpf, ldr, M);
}

void reduce_splits_f32(hipStream_t s, const float *partials, int splits, int64_t n, float *out) {
  if (n % 4 == 0 && ((uintptr_t)partials | (uintptr_t)out) % 16 == 0) {
    int64_t g = (n / 4 + 255) / 256;
    if (g > 8192) g = 8192;
    k_reduce_splits_f32x4<<<(unsigned)g, 256, 0, s>>>((const f32x4 *)partials, splits, n / 4, (f32x4 *)out);
    return;
  }
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  k_reduce_splits_f32<<<(unsigned)g, 256, 0, s>>>(partials, splits, n, out);
}

}  // namespace ptn
