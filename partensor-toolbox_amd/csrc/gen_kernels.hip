// Synthetic data on the device: counter-based Philox4x32-10 draws, X = [[A]],
// Gaussian noise at a given SNR, norms, the explicit reconstruction cost and
// layout conversions.  Draw conventions (shared with the CPU checker):
// counter = (idx_lo, idx_hi, stream, 0), key = (seed_lo, seed_hi); uniform =
// top 53 bits of (w1:w0); normal = Box-Muller cosine branch on (w1:w0),(w3:w2).
// The reference draws from Eigen's Random seeded by time(NULL)
// (DataGeneration.hpp:54-63), which cannot be reproduced; only the X = [[A]]
// semantics (DataGeneration.hpp:120-128) carry over.
#include "ptn_internal.hpp"

#include <cmath>

namespace ptn {

namespace {

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    if (round > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
}

__device__ __forceinline__ void draw4(uint64_t seed, uint32_t stream, uint64_t idx, uint32_t w[4]) {
  w[0] = (uint32_t)idx; w[1] = (uint32_t)(idx >> 32); w[2] = stream; w[3] = 0u;
  philox4x32_10(w, (uint32_t)seed, (uint32_t)(seed >> 32));
}

__device__ __forceinline__ double u53(uint32_t lo, uint32_t hi) {
  uint64_t v = ((((uint64_t)hi) << 32) | (uint64_t)lo) >> 11;
  return (double)v * 0x1.0p-53;
}

__device__ __forceinline__ double draw_uniform(uint64_t seed, uint32_t stream, uint64_t idx) {
  uint32_t w[4];
  draw4(seed, stream, idx, w);
  return u53(w[0], w[1]);
}

__device__ __forceinline__ double draw_normal(uint64_t seed, uint64_t idx) {
  uint32_t w[4];
  draw4(seed, 1u, idx, w);
  double u1 = u53(w[0], w[1]), u2 = u53(w[2], w[3]);
  return sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
}

struct Dims {
  int N;
  int64_t d[kMaxOrder];
  int64_t off[kMaxOrder];  // offsets of the factors in a packed column-major buffer
};

Dims make_dims(int N, const int64_t *dims, int R) {
  Dims D{};
  D.N = N;
  int64_t o = 0;
  for (int k = 0; k < N; ++k) { D.d[k] = dims[k]; D.off[k] = o; o += dims[k] * R; }
  return D;
}

__global__ void k_gen_factors(int64_t n, uint64_t seed, int dist, double *out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    double u = draw_uniform(seed, 0u, (uint64_t)e);
    out[e] = dist == 1 ? -log(1.0 - u) : u;
  }
}

// X(i) = sum_r prod_k A_k(i_k, r): product left to right, then added (no FMA contraction,
// the whole library is built with -ffp-contract=off).
__global__ void k_cp_full(Dims D, int R, int64_t total, const double *__restrict__ pk, double *x64, float *x32) {
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t ik[kMaxOrder];
    int64_t rem = idx;
    for (int k = 0; k < D.N; ++k) { ik[k] = rem % D.d[k]; rem /= D.d[k]; }
    double acc = 0.0;
    for (int r = 0; r < R; ++r) {
      double p = pk[D.off[0] + ik[0] + D.d[0] * r];
      for (int k = 1; k < D.N; ++k) p *= pk[D.off[k] + ik[k] + D.d[k] * r];
      acc += p;
    }
    if (x64) x64[idx] = acc;
    if (x32) x32[idx] = (float)acc;
  }
}

// Block partial sums of squares with a fixed grid-stride order (deterministic).
template <typename F>
__device__ void block_sum_store(double v, double *partials) {
  __shared__ double red[1024 / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
    partials[blockIdx.x] = s;
  }
}

__global__ void k_noise_norm(int64_t n, uint64_t seed, double *partials) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double e = draw_normal(seed, (uint64_t)i);
    s = fma(e, e, s);
  }
  block_sum_store<void>(s, partials);
}

__global__ void k_sqnorm(int64_t n, const double *x64, const float *x32, double *partials) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double v = x64 ? x64[i] : (double)x32[i];
    s = fma(v, v, s);
  }
  block_sum_store<void>(s, partials);
}

__global__ void k_sum_partials(const double *partials, int nparts, double *out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += partials[i];
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < (int)blockDim.x; ++i) t += red[i];
    *out = t;
  }
}

__global__ void k_noise_scale(const double *nx2, const double *ne2, double snr_db, double *scale) {
  *scale = sqrt(*nx2) * pow(10.0, -snr_db / 20.0) / sqrt(*ne2);
}

__global__ void k_add_noise(int64_t n, uint64_t seed, const double *scale, double *x64, float *x32) {
  double s = *scale;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double e = draw_normal(seed, (uint64_t)i);
    if (x64) {
      double v = fma(s, e, x64[i]);
      x64[i] = v;
      if (x32) x32[i] = (float)v;
    } else {
      x32[i] = (float)fma(s, e, (double)x32[i]);
    }
  }
}

__global__ void k_cast(int64_t n, const double *src, float *dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

struct FacPtrs {
  const double *p[kMaxOrder];
};

// Sum over slabs of (X - A_last KR^T)^2 (Cpd.hpp:131-135): reconstruct each element from
// the factor rows, FP64.
__global__ void k_cost_explicit(Dims D, int R, int ldr, FacPtrs F, int64_t total, const double *x64,
                                const float *x32, double *partials) {
  double s = 0.0;
  const int last = D.N - 1;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t ik[kMaxOrder];
    int64_t rem = idx;
    for (int k = 0; k < D.N; ++k) { ik[k] = rem % D.d[k]; rem /= D.d[k]; }
    double rec = 0.0;
    for (int r = 0; r < R; ++r) {
      double p = F.p[0][ik[0] * ldr + r];
      for (int k = 1; k < last; ++k) p *= F.p[k][ik[k] * ldr + r];
      rec = fma(F.p[last][ik[last] * ldr + r], p, rec);
    }
    double d = (x64 ? x64[idx] : (double)x32[idx]) - rec;
    s = fma(d, d, s);
  }
  block_sum_store<void>(s, partials);
}

// Zero fill by vector stores (used where a buffer must start at zero in code a caller may
// capture into a graph, instead of a memset node).
__global__ void k_fill_zero(int64_t n, unsigned long long *p) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    p[e] = 0ull;
}

__global__ void k_colmajor_to_rows(int64_t I, int R, int ldr, const double *src, double *dst) {
  int64_t n = I * ldr;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = e / ldr;
    int r = (int)(e % ldr);
    dst[e] = r < R ? src[i + I * r] : 0.0;
  }
}

__global__ void k_rows_to_colmajor(int64_t I, int R, int ldr, const double *src, double *dst) {
  int64_t n = I * R;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = e % I;
    int64_t r = e / I;
    dst[e] = src[i * ldr + r];
  }
}

int grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

void gen_factors(hipStream_t s, int N, const int64_t *dims, int R, uint64_t seed, int dist, double *packed) {
  int64_t n = 0;
  for (int k = 0; k < N; ++k) n += dims[k] * R;
  k_gen_factors<<<grid_for(n, 256), 256, 0, s>>>(n, seed, dist, packed);
}

void cp_full(hipStream_t s, int N, const int64_t *dims, int R, const double *packed_dev, double *x64, float *x32) {
  Dims D = make_dims(N, dims, R);
  int64_t total = 1;
  for (int k = 0; k < N; ++k) total *= dims[k];
  k_cp_full<<<grid_for(total, 256), 256, 0, s>>>(D, R, total, packed_dev, x64, x32);
}

void gen_noise_norm(hipStream_t s, int64_t n, uint64_t seed, double *partials, int nparts) {
  k_noise_norm<<<nparts, 256, 0, s>>>(n, seed, partials);
}

void add_noise(hipStream_t s, int64_t n, uint64_t seed, const double *scale_dev, double *x64, float *x32) {
  k_add_noise<<<grid_for(n, 256), 256, 0, s>>>(n, seed, scale_dev, x64, x32);
}

void square_norm_partials(hipStream_t s, int64_t n, const double *x64, const float *x32, double *partials, int nparts) {
  k_sqnorm<<<nparts, 256, 0, s>>>(n, x64, x32, partials);
}

void sum_partials(hipStream_t s, const double *partials, int nparts, double *out) {
  k_sum_partials<<<1, 256, 0, s>>>(partials, nparts, out);
}

void noise_scale(hipStream_t s, const double *nx2, const double *ne2, double snr_db, double *scale) {
  k_noise_scale<<<1, 1, 0, s>>>(nx2, ne2, snr_db, scale);
}

void cast_f64_to_f32(hipStream_t s, int64_t n, const double *src, float *dst) {
  k_cast<<<grid_for(n, 256), 256, 0, s>>>(n, src, dst);
}

void cost_explicit_partials(hipStream_t s, int N, const int64_t *dims, int R, int ldr, const double *const *fac,
                            const double *x64, const float *x32, double *partials, int nparts) {
  Dims D = make_dims(N, dims, R);
  FacPtrs F{};
  for (int k = 0; k < N; ++k) F.p[k] = fac[k];
  int64_t total = 1;
  for (int k = 0; k < N; ++k) total *= dims[k];
  k_cost_explicit<<<nparts, 256, 0, s>>>(D, R, ldr, F, total, x64, x32, partials);
}

void fill_zero(hipStream_t s, void *p, int64_t words) {
  if (words > 0) k_fill_zero<<<grid_for(words, 256), 256, 0, s>>>(words, (unsigned long long *)p);
}

void colmajor_to_rows(hipStream_t s, int64_t I, int R, int ldr, const double *src, double *dst) {
  k_colmajor_to_rows<<<grid_for(I * ldr, 256), 256, 0, s>>>(I, R, ldr, src, dst);
}

void rows_to_colmajor(hipStream_t s, int64_t I, int R, int ldr, const double *src, double *dst) {
  k_rows_to_colmajor<<<grid_for(I * R, 256), 256, 0, s>>>(I, R, ldr, src, dst);
}

void rows_to_f32(hipStream_t s, int64_t n, const double *src, float *dst) {
  k_cast<<<grid_for(n, 256), 256, 0, s>>>(n, src, dst);
}

}  // namespace ptn
