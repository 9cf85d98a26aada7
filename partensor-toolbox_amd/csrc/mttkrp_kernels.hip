// MTTKRP building blocks on FP64 matrix cores (v_mfma_f64_16x16x4_f64).
//
// The reference forms M_n = X_(n) * KR explicitly with Eigen GEMM per block
// (MTTKRP.hpp:60-69, :195-220) or streams blocks host->device for cublasDgemm
// (CUDAMTTKRP.hpp:157-175).  Here X stays resident and is read once per
// contraction; Khatri-Rao rows are formed in registers from factor rows; the
// unfolding X_(n) is index arithmetic, never a copy (Matricization.hpp:101-124).
//
// Two GEMM shapes cover every mode of every tree:
//   gemm_cm:  out[m][r] = sum_k X[m + M*k] W(k)[r]   (contracted index strided;
//             mode-0 MTTKRP, the dimension tree's X x_{N-1} A_{N-1})
//   gemm_ck:  T[c][r]   = sum_a X[a + L*c] W(a)[r]   (contracted index contiguous;
//             X x_0 A_0, the dimension tree's right node)
// and contract_rows finishes a contraction against Khatri-Rao rows (the TTVs of
// TensorPartialProduct_R, TensorOperations.hpp:196-213).
//
// MFMA 16x16x4 f64 operand layout (lane l, g = l/16, f = l%16):
//   A (16x4): A[f][g]   B (4x16): B[g][f]   D (16x16): D[g+4v][f], v = 0..3
//   (the D row layout was checked on gfx950 with single-entry tensors)
#include "ptn_internal.hpp"

namespace ptn {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

namespace {

struct KrArgs {
  int nm;
  uint32_t dim[kMaxOrder];
  const double *fac[kMaxOrder];
};

KrArgs to_args(const KrSpec &w) {
  KrArgs a{};
  a.nm = w.nm;
  for (int j = 0; j < w.nm; ++j) { a.dim[j] = (uint32_t)w.dim[j]; a.fac[j] = w.fac[j]; }
  return a;
}

template <typename TX>
__device__ __forceinline__ double ld(const TX *p) { return (double)__builtin_nontemporal_load(p); }

// ---------------------------------------------------------------------------
// gemm_cm: a block owns 64 consecutive rows and one k-chunk; its 4 waves take
// interleaved k-steps of 4 columns and are summed through LDS at the end.
// ---------------------------------------------------------------------------
template <typename TX, int NT>
__global__ __launch_bounds__(256, 2) void k_gemm_cm_f64(const TX *__restrict__ X, int64_t Mrows, int64_t Ktot,
                                                     KrArgs w, int ldr, int64_t kchunk, double *__restrict__ out) {
  constexpr int RP = NT * 16;
  __shared__ double red[64 * RP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, f = lane & 15;
  const int64_t mb = (int64_t)blockIdx.x * 64;
  const int64_t kbeg = (int64_t)blockIdx.y * kchunk;
  const int64_t kend = kbeg + kchunk < Ktot ? kbeg + kchunk : Ktot;

  f64x4 acc[4][NT];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[s][t] = f64x4{0.0, 0.0, 0.0, 0.0};

  // Every load is unconditional, so that loads stay in flight across the MFMAs: a row past
  // Mrows reads row Mrows - 1 (its sums are never stored); a column past the chunk reads
  // the chunk's first column and a factor row 0, and both operands of that column are
  // replaced by zero where the MFMAs use them.
  int64_t roff[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int64_t m = mb + 16 * s + f;
    roff[s] = m < Mrows ? m : Mrows - 1;
  }

  // digits of this lane's first column kk = kbeg + 4*wave + g (lowest mode fastest)
  int64_t kk = kbeg + 4 * wave + g;
  uint32_t dig[kMaxOrder];
  {
    uint64_t rem = (uint64_t)kk;
    for (int j = 0; j < w.nm; ++j) { dig[j] = (uint32_t)(rem % w.dim[j]); rem /= w.dim[j]; }
  }
  // The Khatri-Rao values of the lane's current column as two factors, p (the product of
  // all but the last mode's rows, left to right) and q (the last mode's row; 1 for one
  // mode), so that p * q is the left-to-right product and the last multiply can wait
  // until the MFMAs need it; then the odometer advanced by 16.
  auto kr_next = [&](bool kv, double (&p)[NT], double (&q)[NT]) {
    const int last = w.nm - 1;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      double pv = w.fac[0][(int64_t)(kv ? dig[0] : 0u) * ldr + 16 * t + f];
      for (int j = 1; j < last; ++j) pv *= w.fac[j][(int64_t)(kv ? dig[j] : 0u) * ldr + 16 * t + f];
      p[t] = pv;
      q[t] = last > 0 ? w.fac[last][(int64_t)(kv ? dig[last] : 0u) * ldr + 16 * t + f] : 1.0;
    }
    dig[0] += 16;
    for (int j = 0; j + 1 < w.nm && dig[j] >= w.dim[j]; ++j) {
      uint32_t c = dig[j] / w.dim[j];
      dig[j] -= c * w.dim[j];
      dig[j + 1] += c;
    }
  };
  // one step of 16 columns (4 per wave): loads, then (apply) the zeroing and the MFMAs
  auto load_step = [&](bool &kv, double (&a)[4], double (&p)[NT], double (&q)[NT]) {
    kv = kk < kend;
    kr_next(kv, p, q);
    const TX *xp = X + Mrows * (kv ? kk : kbeg);
#pragma unroll
    for (int s = 0; s < 4; ++s) a[s] = ld(xp + roff[s]);
    kk += 16;
  };
  auto mma_step = [&](bool kv, const double (&a)[4], const double (&p)[NT], const double (&q)[NT]) {
    double b[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) b[t] = kv ? p[t] * q[t] : 0.0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const double av = kv ? a[s] : 0.0;
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b[t], acc[s][t], 0, 0, 0);
    }
  };
  // Groups of KU steps, software-pipelined over two register buffers: group G+1's loads
  // are issued before group G's MFMAs, with no branch between a load and its use inside
  // the loop.  The MFMAs run in step order (the same sums as one step at a time).
  // (FP32 data at two column tiles: groups of one step, since two steps' converted values
  // overflowed the registers into scratch)
  constexpr int KU = NT == 1 ? 4 : NT == 2 && sizeof(TX) == 8 ? 2 : 1;
  struct Buf {
    bool kv[KU];
    double a[KU][4], p[KU][NT], q[KU][NT];
  };
  auto load_group = [&](Buf &B) {
#pragma unroll
    for (int u = 0; u < KU; ++u) load_step(B.kv[u], B.a[u], B.p[u], B.q[u]);
  };
  auto mma_group = [&](const Buf &B) {
#pragma unroll
    for (int u = 0; u < KU; ++u) mma_step(B.kv[u], B.a[u], B.p[u], B.q[u]);
  };
  const int64_t kw = kk - g;                                 // the wave's first column
  const int64_t nsteps = kend > kw ? (kend - kw + 15) / 16 : 0;
  const int64_t ngroups = nsteps / KU;
  if (ngroups > 0) {
    Buf b0, b1;
    load_group(b0);
    int64_t G = 0;
    for (; G + 2 < ngroups; G += 2) {
      load_group(b1);
      mma_group(b0);
      load_group(b0);
      mma_group(b1);
    }
    if (G + 1 < ngroups) {
      load_group(b1);
      mma_group(b0);
      mma_group(b1);
    } else {
      mma_group(b0);
    }
  }
  for (int64_t j = ngroups * KU; j < nsteps; ++j) {
    bool kv;
    double a[4], p[NT], q[NT];
    load_step(kv, a, p, q);
    mma_step(kv, a, p, q);
  }

  // sum the 4 waves' tiles in LDS, in wave order
  for (int src = 0; src < 4; ++src) {
    if (wave == src) {
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            int row = 16 * s + g + 4 * v, col = 16 * t + f;
            red[row * RP + col] = (src == 0 ? 0.0 : red[row * RP + col]) + acc[s][t][v];
          }
    }
    __syncthreads();
  }
  double *o = out + ((int64_t)blockIdx.y * Mrows) * ldr;
  for (int e = threadIdx.x; e < 64 * RP; e += 256) {
    int row = e / RP, col = e % RP;
    if (mb + row < Mrows) o[(mb + row) * ldr + col] = red[e];
  }
}

// ---------------------------------------------------------------------------
// gemm_ck: a wave owns 64 fibers (4 MFMA row tiles) and the whole contraction.
// Within a 16-wide k-block, lane group g supplies a = kb + 4g + t at MFMA step t,
// so each lane reads 4 consecutive elements of its fiber.
// ---------------------------------------------------------------------------
template <typename TX, int NT, bool VEC>
__global__ __launch_bounds__(256, NT <= 2 ? 2 : 1) void k_gemm_ck_f64(const TX *__restrict__ X, int64_t L, int64_t C, KrArgs w,
                                                     int ldr, double *__restrict__ T) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, f = lane & 15;
  const int64_t cb = ((int64_t)blockIdx.x * 4 + wave) * 64;
  if (cb >= C) return;

  f64x4 acc[4][NT];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int u = 0; u < NT; ++u) acc[s][u] = f64x4{0.0, 0.0, 0.0, 0.0};

  bool cv[4];
  const TX *xp[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    int64_t c = cb + 16 * s + f;
    cv[s] = c < C;
    xp[s] = X + L * (cv[s] ? c : 0);
  }
  // the Khatri-Rao row of the contracted modes at a (lowest mode fastest), NT fragments
  // (the product runs over the digits as they are split off, so no array of row pointers
  // is indexed at run time: such an array lives in scratch memory)
  auto kr_row = [&](int64_t a, bool av, double (&b)[NT]) {
#pragma unroll
    for (int u = 0; u < NT; ++u) b[u] = 0.0;
    if (!av) return;
    uint32_t rem = (uint32_t)a;                   // a < L: the last digit is what remains
    for (int j = 0; j < w.nm; ++j) {
      const bool lastd = j + 1 == w.nm;
      const uint32_t d = lastd ? rem : rem % w.dim[j];
      rem = lastd ? 0u : rem / w.dim[j];
      const double *row = w.fac[j] + (int64_t)d * ldr;
#pragma unroll
      for (int u = 0; u < NT; ++u) b[u] = j == 0 ? row[16 * u + f] : b[u] * row[16 * u + f];
    }
  };
  auto mma16 = [&](const double (&x)[4][4], const double (&b)[4][NT]) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int u = 0; u < NT; ++u)
          acc[s][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[s][t], b[t][u], acc[s][u], 0, 0, 0);
  };
  int64_t kb = 0;
  if constexpr (VEC) {
    // L % 4 == 0 and FP64 data: a lane's four elements a = kb + 4g .. kb + 4g + 3 of a fiber
    // are one aligned 32-byte chunk (two 16-byte loads).  Full 16-wide blocks are
    // software-pipelined over two register buffers: block j+1's Khatri-Rao rows and tensor
    // loads are issued before block j's MFMAs.
    const int64_t nblk = L / 16;
    auto load_blk = [&](double (&x)[4][4], double (&b)[4][NT]) {
#pragma unroll
      for (int t = 0; t < 4; ++t) kr_row(kb + 4 * g + t, true, b[t]);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const f64x2 *q = (const f64x2 *)(xp[s] + kb + 4 * g);
        const f64x2 lo = __builtin_nontemporal_load(q), hi = __builtin_nontemporal_load(q + 1);
        x[s][0] = lo[0];                            // a fiber past C reads fiber 0; its
        x[s][1] = lo[1];                            // sums are never stored
        x[s][2] = hi[0];
        x[s][3] = hi[1];
      }
      kb += 16;
    };
    if (nblk > 0) {
      double x0[4][4], b0[4][NT], x1[4][4], b1[4][NT];
      load_blk(x0, b0);
      int64_t j = 0;
      for (; j + 2 < nblk; j += 2) {               // no branch between a load and its use
        load_blk(x1, b1);
        mma16(x0, b0);
        load_blk(x0, b0);
        mma16(x1, b1);
      }
      if (j + 1 < nblk) {
        load_blk(x1, b1);
        mma16(x0, b0);
        mma16(x1, b1);
      } else {
        mma16(x0, b0);
      }
    }
  }
  for (; kb < L; kb += 16) {
    double x[4][4];
    double b[4][NT];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int64_t a = kb + 4 * g + t;
      const bool av = a < L;
#pragma unroll
      for (int s = 0; s < 4; ++s) x[s][t] = (av && cv[s]) ? ld(xp[s] + a) : 0.0;
      kr_row(a, av, b[t]);
    }
    mma16(x, b);
  }
  // the node leaves with non-temporal stores (a leaf reads it after the other streams' work;
  // 800^3 R = 32 right node 1.020 -> 0.993 ms)
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      int64_t c = cb + 16 * s + g + 4 * v;
      if (c < C) {
#pragma unroll
        for (int u = 0; u < NT; ++u) __builtin_nontemporal_store(acc[s][u][v], &T[c * ldr + 16 * u + f]);
      }
    }
}

// ---------------------------------------------------------------------------
// contract_rows: block (i, split) sums, for output index i, the rows of T whose
// remaining digits fall in the split, weighted by the other modes' factor rows.
// Threads: r = tid % RP, group = tid / RP; groups are summed in LDS in order.
// ---------------------------------------------------------------------------
struct RowSpec {
  int nm, outmode;
  uint32_t dim[kMaxOrder];
  int64_t stride[kMaxOrder];   // row stride of each digit in T
  const double *fac[kMaxOrder];
};

template <typename TT>
__global__ __launch_bounds__(256) void k_contract_rows(const TT *__restrict__ T, int64_t ldt, RowSpec sp, int RP,
                                                       int ldr, int64_t cchunk, int64_t Cother,
                                                       double *__restrict__ partials, const int *sel,
                                                       int64_t sel_off) {
  if (sel) T += (int64_t)(*sel) * sel_off;              // node chosen on the device (line search)
  __shared__ double red[256];
  const int r = threadIdx.x % RP, grp = threadIdx.x / RP, ngrp = 256 / RP;
  const uint32_t i = blockIdx.x;
  const int64_t c0 = (int64_t)blockIdx.y * cchunk;
  const int64_t c1 = c0 + cchunk < Cother ? c0 + cchunk : Cother;
  double acc = 0.0;
  if (grp < ngrp) {
    for (int64_t c = c0 + grp; c < c1; c += ngrp) {
      uint64_t rem = (uint64_t)c;
      int64_t row = (int64_t)i * sp.stride[sp.outmode];
      double wv = 1.0;
      for (int j = 0; j < sp.nm; ++j) {
        if (j == sp.outmode) continue;
        uint32_t d = (uint32_t)(rem % sp.dim[j]);
        rem /= sp.dim[j];
        row += (int64_t)d * sp.stride[j];
        wv *= sp.fac[j][(int64_t)d * ldr + r];
      }
      acc = fma((double)T[row * ldt + r], wv, acc);
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < RP) {
    double s = 0.0;
    for (int q = 0; q < ngrp; ++q) s += red[q * RP + threadIdx.x];
    partials[((int64_t)blockIdx.y * gridDim.x + i) * ldr + threadIdx.x] = s;
  }
}

// ---------------------------------------------------------------------------
// contract2: a node with two digits (rows m = i0 + D0 i1, ld values per row) contracted
// over one digit against the other digit's factor rows (a leaf TTV of the dimension
// tree, TensorPartialProduct_R, TensorOperations.hpp:196-213).  Thread (q, c): row
// group q = tid / 32, column c (+ 32 t for the column tiles).  Loads are issued four
// rows at a time; the order of the sums is fixed.
//   outmode 0: block (8 consecutive i0, i1-split), out[i0][r] = sum_i1 Y[m][r] F[i1][r]
//   outmode 1: block (one i1, i0-split), 8 row groups summed in LDS in order
// With gridDim.z == 2 the blocks of z = 1 contract (Y2, F2) beside (Y, F); partials are
// part[split][z][i][ldr].
// ---------------------------------------------------------------------------
template <typename TT>
__global__ __launch_bounds__(256) void k_contract2(const TT *__restrict__ Y, int64_t ld, int64_t D0, int64_t D1,
                                                   int outmode, const double *__restrict__ F, int ldr, int R,
                                                   int64_t slen, double *__restrict__ part, const int *sel,
                                                   int64_t sel_off, const TT *__restrict__ Y2,
                                                   const double *__restrict__ F2) {
  if (sel) Y += (int64_t)(*sel) * sel_off;              // node chosen on the device (line search)
  if (blockIdx.z) {
    Y = Y2;
    F = F2;
  }
  const int64_t zs = (int64_t)blockIdx.y * gridDim.z + blockIdx.z;
  __shared__ double red[8][32];
  const int q = threadIdx.x >> 5, c0 = threadIdx.x & 31;
  const int NT = (R + 31) / 32;
  if (outmode == 0) {
    const int64_t i0 = (int64_t)blockIdx.x * 8 + q;
    const int64_t b = (int64_t)blockIdx.y * slen, e = b + slen < D1 ? b + slen : D1;
    for (int t = 0; t < NT; ++t) {
      const int c = 32 * t + c0;
      if (i0 >= D0 || c >= ldr) continue;
      double acc = 0.0;
      int64_t i1 = b;
      for (; i1 + 4 <= e; i1 += 4) {
        const double y0 = (double)Y[(i0 + D0 * i1) * ld + c], y1 = (double)Y[(i0 + D0 * (i1 + 1)) * ld + c];
        const double y2 = (double)Y[(i0 + D0 * (i1 + 2)) * ld + c], y3 = (double)Y[(i0 + D0 * (i1 + 3)) * ld + c];
        const double f0 = F[i1 * ldr + c], f1 = F[(i1 + 1) * ldr + c];
        const double f2 = F[(i1 + 2) * ldr + c], f3 = F[(i1 + 3) * ldr + c];
        acc = fma(y0, f0, acc);
        acc = fma(y1, f1, acc);
        acc = fma(y2, f2, acc);
        acc = fma(y3, f3, acc);
      }
      for (; i1 < e; ++i1) acc = fma((double)Y[(i0 + D0 * i1) * ld + c], F[i1 * ldr + c], acc);
      part[(zs * D0 + i0) * ldr + c] = acc;
    }
    return;
  }
  const int64_t i1 = blockIdx.x;
  const int64_t b = (int64_t)blockIdx.y * slen, e = b + slen < D0 ? b + slen : D0;
  for (int t = 0; t < NT; ++t) {
    const int c = 32 * t + c0;
    double acc = 0.0;
    if (c < ldr) {
      const TT *row = Y + (D0 * i1) * ld + c;
      int64_t i0 = b + q;
      for (; i0 + 24 < e; i0 += 32) {
        const double y0 = (double)row[i0 * ld], y1 = (double)row[(i0 + 8) * ld];
        const double y2 = (double)row[(i0 + 16) * ld], y3 = (double)row[(i0 + 24) * ld];
        const double f0 = F[i0 * ldr + c], f1 = F[(i0 + 8) * ldr + c];
        const double f2 = F[(i0 + 16) * ldr + c], f3 = F[(i0 + 24) * ldr + c];
        acc = fma(y0, f0, acc);
        acc = fma(y1, f1, acc);
        acc = fma(y2, f2, acc);
        acc = fma(y3, f3, acc);
      }
      for (; i0 < e; i0 += 8) acc = fma((double)row[i0 * ld], F[i0 * ldr + c], acc);
    }
    __syncthreads();
    red[q][c0] = acc;
    __syncthreads();
    if (q == 0 && c < ldr) {
      double sum = red[0][c0];
      for (int g = 1; g < 8; ++g) sum += red[g][c0];
      part[(zs * D1 + i1) * ldr + c] = sum;
    }
  }
}

// k_contract2's outmode 0 for FP32 nodes with 16-byte loads: thread (rr = tid / 8, c4 = tid % 8)
// owns row i0 = 32 bx + rr and columns 4 c4 .. 4 c4 + 3 of each 32-column group; each column
// sums its i1 range in the same order as k_contract2, so the partials are bitwise equal.
// Rows of Y hold ld floats, ld and the node's offset multiples of 4 (16-byte aligned).
__global__ __launch_bounds__(256) void k_contract2_r4(const float *__restrict__ Y, int64_t ld, int64_t D0, int64_t D1,
                                                      const double *__restrict__ F, int ldr, int64_t slen,
                                                      double *__restrict__ part, const int *sel, int64_t sel_off,
                                                      const float *__restrict__ Y2, const double *__restrict__ F2) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  if (sel) Y += (int64_t)(*sel) * sel_off;
  if (blockIdx.z) {
    Y = Y2;
    F = F2;
  }
  const int64_t zs = (int64_t)blockIdx.y * gridDim.z + blockIdx.z;
  const int rr = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
  const int64_t i0 = (int64_t)blockIdx.x * 32 + rr;
  if (i0 >= D0) return;
  const int64_t b = (int64_t)blockIdx.y * slen, e = b + slen < D1 ? b + slen : D1;
  for (int cb = 0; cb < ldr; cb += 32) {
    const int c = cb + c4;
    if (c >= ldr) break;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t i1 = b;
    for (; i1 + 4 <= e; i1 += 4) {
      f4 y[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) y[u] = *(const f4 *)(Y + (i0 + D0 * (i1 + u)) * ld + c);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double *f = F + (i1 + u) * ldr + c;
        a0 = fma((double)y[u][0], f[0], a0);
        a1 = fma((double)y[u][1], f[1], a1);
        a2 = fma((double)y[u][2], f[2], a2);
        a3 = fma((double)y[u][3], f[3], a3);
      }
    }
    for (; i1 < e; ++i1) {
      const f4 y = *(const f4 *)(Y + (i0 + D0 * i1) * ld + c);
      const double *f = F + i1 * ldr + c;
      a0 = fma((double)y[0], f[0], a0);
      a1 = fma((double)y[1], f[1], a1);
      a2 = fma((double)y[2], f[2], a2);
      a3 = fma((double)y[3], f[3], a3);
    }
    double *o = part + (zs * D0 + i0) * ldr + c;
    o[0] = a0;
    o[1] = a1;
    o[2] = a2;
    o[3] = a3;
  }
}

__global__ void k_reduce_splits(const double *__restrict__ partials, int splits, int64_t n, double *__restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int q = 0; q < splits; ++q) s += partials[q * n + e];
    out[e] = s;
  }
}
// the same sums two at a time (n % 2 == 0): 16-byte loads of every split, in split order
__global__ void k_reduce_splits_x2(const f64x2 *__restrict__ partials, int splits, int64_t n2,
                                   f64x2 *__restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n2; e += (int64_t)gridDim.x * blockDim.x) {
    f64x2 s = {0.0, 0.0};
    for (int q = 0; q < splits; ++q) {
      const f64x2 v = partials[q * n2 + e];
      s[0] += v[0];
      s[1] += v[1];
    }
    out[e] = s;
  }
}

}  // namespace

void gemm_cm_f64(hipStream_t s, const double *x64, const float *x32, int64_t Mrows, int64_t Ktot, const KrSpec &w,
                 int ldr, int splits, double *out) {
  KrArgs a = to_args(w);
  int64_t kchunk = (Ktot + splits - 1) / splits;
  dim3 grid((unsigned)((Mrows + 63) / 64), (unsigned)splits);
  const int NT = ldr / 16;
#define PTN_CM(NTV)                                                                              \
  if (NT == NTV) {                                                                               \
    if (x64) k_gemm_cm_f64<double, NTV><<<grid, 256, 0, s>>>(x64, Mrows, Ktot, a, ldr, kchunk, out); \
    else     k_gemm_cm_f64<float, NTV><<<grid, 256, 0, s>>>(x32, Mrows, Ktot, a, ldr, kchunk, out);  \
    return;                                                                                      \
  }
  PTN_CM(1) PTN_CM(2) PTN_CM(3) PTN_CM(4)
#undef PTN_CM
}

void gemm_ck_f64(hipStream_t s, const double *x64, const float *x32, int64_t L, int64_t C, const KrSpec &w, int ldr,
                 double *T) {
  KrArgs a = to_args(w);
  dim3 grid((unsigned)((C + 255) / 256));
  const int NT = ldr / 16;
  const bool vec = x64 && L % 4 == 0;         // aligned 32-byte chunks of every fiber
#define PTN_CK(NTV)                                                                          \
  if (NT == NTV) {                                                                           \
    if (vec) k_gemm_ck_f64<double, NTV, true><<<grid, 256, 0, s>>>(x64, L, C, a, ldr, T);     \
    else if (x64) k_gemm_ck_f64<double, NTV, false><<<grid, 256, 0, s>>>(x64, L, C, a, ldr, T); \
    else     k_gemm_ck_f64<float, NTV, false><<<grid, 256, 0, s>>>(x32, L, C, a, ldr, T);     \
    return;                                                                                  \
  }
  PTN_CK(1) PTN_CK(2) PTN_CK(3) PTN_CK(4)
#undef PTN_CK
}

void contract_rows(hipStream_t s, const double *T64, const float *T32, int64_t ldt, const KrSpec &spec, int outmode,
                   int ldr, int splits, double *partials, const int *sel, int64_t sel_off) {
  RowSpec sp{};
  sp.nm = spec.nm;
  sp.outmode = outmode;
  int64_t st = 1, Cother = 1;
  for (int j = 0; j < spec.nm; ++j) {
    sp.dim[j] = (uint32_t)spec.dim[j];
    sp.stride[j] = st;
    st *= spec.dim[j];
    sp.fac[j] = spec.fac[j];
    if (j != outmode) Cother *= spec.dim[j];
  }
  int RP = ldr;
  int64_t cchunk = (Cother + splits - 1) / splits;
  dim3 grid((unsigned)spec.dim[outmode], (unsigned)splits);
  if (T64) k_contract_rows<double><<<grid, 256, 0, s>>>(T64, ldt, sp, RP, ldr, cchunk, Cother, partials, sel, sel_off);
  else     k_contract_rows<float><<<grid, 256, 0, s>>>(T32, ldt, sp, RP, ldr, cchunk, Cother, partials, sel, sel_off);
}

void contract2(hipStream_t s, const double *Y64, const float *Y32, int64_t ld, int64_t D0, int64_t D1, int outmode,
               const double *F, int ldr, int R, int splits, double *part, const int *sel, int64_t sel_off) {
  const int64_t D = outmode == 0 ? D1 : D0;             // contracted extent
  const int64_t slen = (D + splits - 1) / splits;
  dim3 grid(outmode == 0 ? (unsigned)((D0 + 7) / 8) : (unsigned)D1, (unsigned)splits);
  if (Y64)
    k_contract2<double><<<grid, 256, 0, s>>>(Y64, ld, D0, D1, outmode, F, ldr, R, slen, part, sel, sel_off, nullptr,
                                             nullptr);
  else if (outmode == 0 && ld % 4 == 0 && sel_off % 4 == 0 && ((uintptr_t)Y32 & 15) == 0)
    k_contract2_r4<<<dim3((unsigned)((D0 + 31) / 32), (unsigned)splits), 256, 0, s>>>(
        Y32, ld, D0, D1, F, ldr, slen, part, sel, sel_off, nullptr, nullptr);
  else
    k_contract2<float><<<grid, 256, 0, s>>>(Y32, ld, D0, D1, outmode, F, ldr, R, slen, part, sel, sel_off, nullptr,
                                            nullptr);
}

void contract2_pair(hipStream_t s, const float *Ya, const float *Yb, int64_t ld, int64_t D0, int64_t D1, int outmode,
                    const double *Fa, const double *Fb, int ldr, int R, int splits, double *part) {
  const int64_t D = outmode == 0 ? D1 : D0;
  const int64_t slen = (D + splits - 1) / splits;
  if (outmode == 0 && ld % 4 == 0 && ((uintptr_t)Ya & 15) == 0 && ((uintptr_t)Yb & 15) == 0) {
    k_contract2_r4<<<dim3((unsigned)((D0 + 31) / 32), (unsigned)splits, 2), 256, 0, s>>>(
        Ya, ld, D0, D1, Fa, ldr, slen, part, nullptr, 0, Yb, Fb);
    return;
  }
  dim3 grid(outmode == 0 ? (unsigned)((D0 + 7) / 8) : (unsigned)D1, (unsigned)splits, 2);
  k_contract2<float><<<grid, 256, 0, s>>>(Ya, ld, D0, D1, outmode, Fa, ldr, R, slen, part, nullptr, 0, Yb, Fb);
}

void reduce_splits(hipStream_t s, const double *partials, int splits, int64_t n, double *out) {
  if (n % 2 == 0 && ((uintptr_t)partials | (uintptr_t)out) % 16 == 0) {
    int64_t g = (n / 2 + 255) / 256;
    if (g > 8192) g = 8192;
    k_reduce_splits_x2<<<(unsigned)g, 256, 0, s>>>((const f64x2 *)partials, splits, n / 2, (f64x2 *)out);
    return;
  }
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  k_reduce_splits<<<(unsigned)g, 256, 0, s>>>(partials, splits, n, out);
}

}  // namespace ptn
