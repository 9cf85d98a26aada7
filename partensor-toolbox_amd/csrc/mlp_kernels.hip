// The learned NNLS warm start: a 3-layer MLP (2208 -> 256 -> 256 -> 64, ReLU) that maps
// one factor row's subproblem (the Gram V, the MTTKRP row, the previous row) to a start
// point for NesterovMNLS.  BF16 weights and activations, FP32 accumulation on
// v_mfma_f32_32x32x16_bf16, FP32 master weights updated by Adam.  The subproblem itself is
// never changed: the network only proposes the start, the proximal anchor stays the
// previous factor row, and a row keeps the previous row as its start unless the proposal
// passes the KKT test in k_warm_select.
#include "ptn_internal.hpp"

namespace ptn {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// C[m][n] = sum_k A[m][k] B[n][k] (+ bias[n]) (* (gate[m][n] > 0)) (ReLU), both operands
// bf16 with k contiguous (K a multiple of 16, leading dimensions multiples of 8).  Each
// wave owns a 32 x 32 tile; lane (i, kg) loads the 16-byte k-slices A[m0+i][k+8kg..] and
// B[n0+i][k+8kg..] straight into its MFMA fragments.  Rows / columns past M / N read a
// clamped row and are not stored.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gemm_nt(int M, int N, int K, const __bf16 *__restrict__ A, int lda,
                                                 const __bf16 *__restrict__ B, int ldb, const float *__restrict__ bias,
                                                 const __bf16 *__restrict__ gate, int ldg, int relu,
                                                 float *__restrict__ C, int ldc, __bf16 *__restrict__ Cb, int ldcb) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = lane & 31, kg = lane >> 5;
  const int m0 = blockIdx.y * 64 + 32 * (wave >> 1), n0 = blockIdx.x * 64 + 32 * (wave & 1);
  if (m0 >= M || n0 >= N) return;                       // no block-wide barrier below
  const int ma = m0 + i < M ? m0 + i : M - 1, nb = n0 + i < N ? n0 + i : N - 1;
  const bf16x8 *pa = (const bf16x8 *)(A + (int64_t)ma * lda + 8 * kg);
  const bf16x8 *pb = (const bf16x8 *)(B + (int64_t)nb * ldb + 8 * kg);
  f32x16 acc;
#pragma unroll
  for (int v = 0; v < 16; ++v) acc[v] = 0.f;
  int k = 0;
  for (; k + 32 <= K; k += 32) {
    const bf16x8 a0 = pa[k / 8], b0 = pb[k / 8], a1 = pa[k / 8 + 2], b1 = pb[k / 8 + 2];
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc, 0, 0, 0);
  }
  for (; k < K; k += 16) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pa[k / 8], pb[k / 8], acc, 0, 0, 0);
  const int n = n0 + i;
  if (n >= N) return;
  const float bn = bias ? bias[n] : 0.f;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int m = m0 + 8 * (v / 4) + 4 * kg + (v % 4);
    if (m < M) {
      float x = acc[v] + bn;
      if (gate && !((float)gate[(int64_t)m * ldg + n] > 0.f)) x = 0.f;
      if (relu) x = x > 0.f ? x : 0.f;
      if (C) C[(int64_t)m * ldc + n] = x;
      if (Cb) Cb[(int64_t)m * ldcb + n] = (__bf16)x;
    }
  }
}

// out[c][r] = in[r][c] (bf16), 32 x 32 tiles through LDS
__global__ __launch_bounds__(256) void k_transpose_bf16(int rows, int cols, const __bf16 *__restrict__ in, int ldi,
                                                        __bf16 *__restrict__ out, int ldo) {
  __shared__ __bf16 t[32][33];
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int y = ty; y < 32; y += 8) {
    const int r = r0 + y, c = c0 + tx;
    t[y][tx] = (r < rows && c < cols) ? in[(int64_t)r * ldi + c] : (__bf16)0.f;
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const int c = c0 + y, r = r0 + tx;
    if (c < cols && r < rows) out[(int64_t)c * ldo + r] = t[tx][y];
  }
}

// ---------------------------------------------------------------------------
// Features of factor rows (one 64-lane wave per row, lane = column c < 64):
//   [ vech(V) / sV (upper triangle of the 64 x 64 zero-padded V, row by row),
//     m / s_m, a_prev * sV / s_m ],   sV = max_k V_kk, s_m = max_c |m_c| (1 if 0).
// The target of a solution row a* is a* sV / s_m.  scale[row] = s_m / sV maps the
// network's output back.  With Anew: targets T[row][c] (zero beyond R) and rank[row] = R.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int vech_index(int a, int b) {       // a <= b < 64
  return a * kMlpR - (a * (a - 1)) / 2 + (b - a);
}

__global__ __launch_bounds__(256) void k_mlp_features(int64_t I, int R, int ldr, const double *__restrict__ V,
                                                      const double *__restrict__ M, const double *__restrict__ A,
                                                      const double *__restrict__ Anew, __bf16 *__restrict__ X,
                                                      float *__restrict__ T, unsigned char *__restrict__ rank,
                                                      float *__restrict__ scale) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= I) return;
  double sv = 0.0;
  for (int k = 0; k < R; ++k) sv = fmax(sv, V[k * R + k]);
  if (!(sv > 0.0)) sv = 1.0;
  const bool on = lane < R;
  const double m = on ? M[row * ldr + lane] : 0.0;
  double sm = fabs(m);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) sm = fmax(sm, __shfl_xor(sm, o, 64));
  if (!(sm > 0.0)) sm = 1.0;
  __bf16 *x = X + row * kMlpIn;
  // lane c writes row c of the upper triangle: entries (c, c..63)
  for (int b = lane; b < kMlpR; ++b) {
    const double v = (lane < R && b < R) ? V[lane * R + b] / sv : 0.0;
    x[vech_index(lane, b)] = (__bf16)(float)v;
  }
  x[kMlpVech + lane] = (__bf16)(float)(m / sm);
  x[kMlpVech + kMlpR + lane] = (__bf16)(float)(on ? A[row * ldr + lane] * sv / sm : 0.0);
  if (scale && lane == 0) scale[row] = (float)(sm / sv);
  if (T) T[row * kMlpR + lane] = (float)(on ? Anew[row * ldr + lane] * sv / sm : 0.0);
  if (rank && lane == 0) rank[row] = (unsigned char)R;
}

// ---------------------------------------------------------------------------
// The warm-start decision (one wave per row, lane = column).  With Zt = V + lambda I,
// wt = m + lambda a_prev (the subproblem NesterovMNLS solves, anchor unchanged) and L its
// Lipschitz constant, the relative KKT residual of a start a is the projected-gradient
// fixed-point residual min(a, g / L), g = a Zt - wt, relative to the solution's scale
// max|wt| / L:  r(a) = max_c |min(L a_c, g_c)| / max_c |wt_c|.
// The row starts from the network's proposal a_p = max(y, 0) * scale when y is finite and
// r(a_p) <= kkt_tol; otherwise from a_prev (the fallback, which is exactly the reference's
// start).  With C, the proposal is C's row instead (max(C, 0): the proposals after some
// Nesterov steps, whose count, prm->iters, is kept in prm->warm_steps).  S receives the
// start rows (S may be C); count, when given, the rows that took the proposal.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_warm_select(int64_t I, int R, int ldr, const double *__restrict__ V,
                                                     NnlsParams *prm, const double *__restrict__ M,
                                                     const double *__restrict__ A, const float *__restrict__ Y,
                                                     const float *__restrict__ scale, const double *C,
                                                     double kkt_tol, double *S, int *count) {
  if (C && blockIdx.x == 0 && threadIdx.x == 0) prm->warm_steps = prm->iters;
  __shared__ double buf[4][kMlpR];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t row = (int64_t)blockIdx.x * 4 + wave;
  if (row >= I) return;
  const bool on = lane < R;
  const double lambda = prm->lambda, L = prm->L;
  const double ap = on ? A[row * ldr + lane] : 0.0;
  bool fin;
  double aw;
  if (C) {
    const double c = on ? C[row * ldr + lane] : 0.0;
    fin = !on || isfinite(c);
    aw = c > 0.0 ? c : 0.0;
  } else {
    const float y = on ? Y[row * kMlpR + lane] : 0.f;
    fin = !on || isfinite(y);
    aw = (on && y > 0.f) ? (double)y * (double)scale[row] : 0.0;
  }
  buf[wave][lane] = aw;
  __builtin_amdgcn_wave_barrier();
  double gw = 0.0;
  if (on) {
    for (int k = 0; k < R; ++k) {
      const double z = V[k * R + lane] + (k == lane ? lambda : 0.0);
      gw = fma(buf[wave][k], z, gw);
    }
  }
  const double wt = on ? M[row * ldr + lane] + lambda * ap : 0.0;
  gw -= wt;
  double rw = fabs(fmin(L * aw, gw)), sw = fabs(wt);
  if (!on) rw = sw = 0.0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    rw = fmax(rw, __shfl_xor(rw, o, 64));
    sw = fmax(sw, __shfl_xor(sw, o, 64));
  }
  const bool all_fin = __ballot(!fin) == 0ull;
  const double den = sw > 0.0 ? sw : 1.0;
  const bool take = all_fin && rw / den <= kkt_tol;
  if (lane < ldr) S[row * ldr + lane] = take ? (on ? aw : 0.0) : (lane < R ? ap : A[row * ldr + lane]);
  if (count && take && lane == 0) atomicAdd(count, 1);
}

// loss = sum over valid entries (c < rank[row]) of (Y - T)^2 / count; dY = 2 (Y - T) / count
// (bf16 for the weight gradient, FP32 for the bias gradient); partial sums per block.
__global__ __launch_bounds__(256) void k_mse_grad(int B, const float *__restrict__ Y, const float *__restrict__ T,
                                                  const unsigned char *__restrict__ rank, float inv_count,
                                                  __bf16 *__restrict__ dYb, float *__restrict__ dY,
                                                  double *__restrict__ part) {
  __shared__ double red[256];
  const int64_t n = (int64_t)B * kMlpR;
  double s = 0.0;
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const int64_t r = e / kMlpR;
    const int c = (int)(e - r * kMlpR);
    float g = 0.f;
    if (c < rank[r]) {
      const float d = Y[e] - T[e];
      s += (double)d * d;
      g = 2.f * d * inv_count;
    }
    dYb[e] = (__bf16)g;
    dY[e] = g;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] * inv_count;
}

// bias gradient: out[n] = sum_m G[m][n] (FP32 or bf16 rows), one thread per column, rows in order
template <typename TG>
__global__ void k_colsum(int M, int N, const TG *__restrict__ G, int ldg, float *__restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += (float)G[(int64_t)m * ldg + n];
  out[n] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 master weights; the bf16 copies of the three weight
// matrices follow (the biases are used in FP32).
__global__ void k_adam(int64_t n, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2,
                       __bf16 *__restrict__ wb) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const float gr = g[e];
    const float a = b1 * m1[e] + (1.f - b1) * gr;
    const float b = b2 * m2[e] + (1.f - b2) * gr * gr;
    m1[e] = a;
    m2[e] = b;
    const float x = w[e] - lr * (a * c1) / (sqrtf(b * c2) + eps);
    w[e] = x;
    const int64_t bi = mlp_bf16_index(e);
    if (bi >= 0) wb[bi] = (__bf16)x;
  }
}

// bf16 copies of the weight matrices from the FP32 master vector
__global__ void k_weights_bf16(int64_t n, const float *__restrict__ w, __bf16 *__restrict__ wb) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t bi = mlp_bf16_index(e);
    if (bi >= 0) wb[bi] = (__bf16)w[e];
  }
}

// rows idx[0..B) of a data set into contiguous batch buffers (X bf16, T FP32, rank)
__global__ __launch_bounds__(256) void k_gather(int B, const int *__restrict__ idx, const __bf16 *__restrict__ X,
                                                const float *__restrict__ T, const unsigned char *__restrict__ rank,
                                                __bf16 *__restrict__ Xb, float *__restrict__ Tb,
                                                unsigned char *__restrict__ rb) {
  const int r = blockIdx.x;
  if (r >= B) return;
  const int64_t src = idx[r];
  const uint4 *xs = (const uint4 *)(X + src * kMlpIn);
  uint4 *xd = (uint4 *)(Xb + (int64_t)r * kMlpIn);
  for (int e = threadIdx.x; e < kMlpIn / 8; e += blockDim.x) xd[e] = xs[e];
  if (threadIdx.x < kMlpR) Tb[(int64_t)r * kMlpR + threadIdx.x] = T[src * kMlpR + threadIdx.x];
  if (threadIdx.x == 0) rb[r] = rank[src];
}

}  // namespace

void gemm_nt(hipStream_t s, int M, int N, int K, const void *A, int lda, const void *B, int ldb, const float *bias,
             const void *gate, int ldg, int relu, float *C, int ldc, void *Cb, int ldcb) {
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  k_gemm_nt<<<grid, 256, 0, s>>>(M, N, K, (const __bf16 *)A, lda, (const __bf16 *)B, ldb, bias, (const __bf16 *)gate,
                                 ldg, relu, C, ldc, (__bf16 *)Cb, ldcb);
}

void transpose_bf16(hipStream_t s, int rows, int cols, const void *in, int ldi, void *out, int ldo) {
  dim3 grid((cols + 31) / 32, (rows + 31) / 32);
  k_transpose_bf16<<<grid, 256, 0, s>>>(rows, cols, (const __bf16 *)in, ldi, (__bf16 *)out, ldo);
}

void mlp_features(hipStream_t s, int64_t I, int R, int ldr, const double *V, const double *M, const double *A,
                  const double *Anew, void *X, float *T, unsigned char *rank, float *scale) {
  k_mlp_features<<<(unsigned)((I + 3) / 4), 256, 0, s>>>(I, R, ldr, V, M, A, Anew, (__bf16 *)X, T, rank, scale);
}

void warm_select(hipStream_t s, int64_t I, int R, int ldr, const double *V, NnlsParams *p, const double *M,
                 const double *A, const float *Y, const float *scale, const double *C, double kkt_tol, double *S,
                 int *count) {
  k_warm_select<<<(unsigned)((I + 3) / 4), 256, 0, s>>>(I, R, ldr, V, p, M, A, Y, scale, C, kkt_tol, S, count);
}

void mse_grad(hipStream_t s, int B, const float *Y, const float *T, const unsigned char *rank, float inv_count,
              void *dYb, float *dY, double *part, int nparts) {
  k_mse_grad<<<nparts, 256, 0, s>>>(B, Y, T, rank, inv_count, (__bf16 *)dYb, dY, part);
}

void colsum_f32(hipStream_t s, int M, int N, const float *G, int ldg, float *out) {
  k_colsum<float><<<(N + 255) / 256, 256, 0, s>>>(M, N, G, ldg, out);
}

void colsum_bf16(hipStream_t s, int M, int N, const void *G, int ldg, float *out) {
  k_colsum<__bf16><<<(N + 255) / 256, 256, 0, s>>>(M, N, (const __bf16 *)G, ldg, out);
}

void adam_step(hipStream_t s, int64_t n, float *w, const float *g, float *m1, float *m2, float lr, float b1, float b2,
               float eps, int step, void *wb) {
  const float c1 = 1.f / (1.f - powf(b1, (float)step)), c2 = 1.f / (1.f - powf(b2, (float)step));
  k_adam<<<1024, 256, 0, s>>>(n, w, g, m1, m2, lr, b1, b2, eps, c1, c2, (__bf16 *)wb);
}

void weights_bf16(hipStream_t s, int64_t n, const float *w, void *wb) {
  k_weights_bf16<<<1024, 256, 0, s>>>(n, w, (__bf16 *)wb);
}

void gather_rows(hipStream_t s, int B, const int *idx, const void *X, const float *T, const unsigned char *rank,
                 void *Xb, float *Tb, unsigned char *rb) {
  k_gather<<<B, 256, 0, s>>>(B, idx, (const __bf16 *)X, T, rank, (__bf16 *)Xb, Tb, rb);
}

}  // namespace ptn
