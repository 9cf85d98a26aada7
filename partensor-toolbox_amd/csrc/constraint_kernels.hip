// The other branches of update_factor (Cpd.hpp:189-228) on the device:
//   unconstrained   A = M V^{-1}         (Cpd.hpp:196-200; Eigen .inverse(), PartialPivLU)
//   orthogonality   A = M (M^T M)^{-1/2}  (Cpd.hpp:210-219; symmetric eigendecomposition)
// R x R work runs in one block with the matrices in LDS; every element is computed with
// the operations, in the order, of the CPU restatement (ptno_inverse, jacobi_rr,
// update_factor), so the results agree bit for bit.  Matrices are column-major R x R
// (a[i + R j]); factor rows are row-major with leading dimension ldr.
#include "ptn_internal.hpp"

namespace ptn {

namespace {

constexpr double kDblEps = 2.220446049250313e-16;

// LU with partial pivoting (the largest |entry| of the column, first one on ties), then
// the inverse column by column: pivots, forward substitution (unit lower), backward
// substitution.  A zero pivot leaves Vi as NaN.
__global__ __launch_bounds__(256) void k_lu_inverse(int R, const double *__restrict__ V, double *__restrict__ Vi) {
  __shared__ double lu[kMaxRank * kMaxRank], x[kMaxRank * kMaxRank];
  __shared__ int piv[kMaxRank];
  __shared__ int ok;
  const int t = threadIdx.x;
  for (int e = t; e < R * R; e += 256) lu[e] = V[e];
  if (t == 0) ok = 1;
  __syncthreads();
  for (int k = 0; k < R; ++k) {
    if (t == 0) {
      int p = k;
      for (int i = k + 1; i < R; ++i)
        if (fabs(lu[i + R * k]) > fabs(lu[p + R * k])) p = i;
      piv[k] = p;
      if (lu[p + R * k] == 0.0) ok = 0;
    }
    __syncthreads();
    if (!ok) break;
    const int p = piv[k];
    if (p != k)
      for (int j = t; j < R; j += 256) {
        const double tmp = lu[k + R * j];
        lu[k + R * j] = lu[p + R * j];
        lu[p + R * j] = tmp;
      }
    __syncthreads();
    for (int i = k + 1 + t; i < R; i += 256) lu[i + R * k] = lu[i + R * k] / lu[k + R * k];
    __syncthreads();
    const int m = R - k - 1;
    for (int e = t; e < m * m; e += 256) {
      const int i = k + 1 + e % m, j = k + 1 + e / m;
      lu[i + R * j] = fma(-lu[i + R * k], lu[k + R * j], lu[i + R * j]);
    }
    __syncthreads();
  }
  if (!ok) {
    for (int e = t; e < R * R; e += 256) Vi[e] = NAN;
    return;
  }
  if (t < R) {
    const int c = t;
    double *xc = x + R * c;
    for (int i = 0; i < R; ++i) xc[i] = (i == c) ? 1.0 : 0.0;
    for (int k = 0; k < R; ++k)
      if (piv[k] != k) {
        const double tmp = xc[k];
        xc[k] = xc[piv[k]];
        xc[piv[k]] = tmp;
      }
    for (int i = 0; i < R; ++i)
      for (int k = 0; k < i; ++k) xc[i] = fma(-lu[i + R * k], xc[k], xc[i]);
    for (int i = R - 1; i >= 0; --i) {
      for (int k = i + 1; k < R; ++k) xc[i] = fma(-lu[i + R * k], xc[k], xc[i]);
      xc[i] = xc[i] / lu[i + R * i];
    }
  }
  __syncthreads();
  for (int e = t; e < R * R; e += 256) Vi[e] = x[e];
}

__device__ __forceinline__ int rr_slot(int k, int t, int n) { return k == 0 ? 0 : ((k - 1 + t) % (n - 1)) + 1; }

// Parallel-order cyclic Jacobi (Golub & Van Loan 8.5, Brent-Luk round-robin order): each
// round applies the n/2 disjoint rotations, rows first (J^T A), then columns (A J) and the
// eigenvector accumulation; a rotation is skipped when |a_pq| <= eps sqrt(|a_pp a_qq|).
// Up to 60 sweeps, stopping after a sweep without rotations.  evals = diag, Q = vectors.
__global__ __launch_bounds__(256) void k_jacobi_eig(int R, const double *__restrict__ A, double *__restrict__ evals,
                                                    double *__restrict__ Q) {
  __shared__ double a[kMaxRank * kMaxRank], v[kMaxRank * kMaxRank];
  __shared__ double cs[kMaxRank / 2], sn[kMaxRank / 2];
  __shared__ int pp[kMaxRank / 2], qq[kMaxRank / 2];
  __shared__ int np, rotated;
  const int t = threadIdx.x;
  const int n = R + (R & 1);
  for (int e = t; e < R * R; e += 256) {
    a[e] = A[e];
    v[e] = (e % R == e / R) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 60; ++sweep) {
    if (t == 0) rotated = 0;
    for (int r = 0; r < n - 1; ++r) {
      __syncthreads();                               // everyone has read the last round's np
      if (t == 0) {
        // the rotations of this round, in pair order (they touch disjoint rows / columns)
        int c = 0;
        for (int k = 0; k < n / 2; ++k) {
          int p = rr_slot(k, r, n), q = rr_slot(n - 1 - k, r, n);
          if (p > q) {
            const int tmp = p;
            p = q;
            q = tmp;
          }
          if (q >= R) continue;
          const double apq = a[p + R * q], app = a[p + R * p], aqq = a[q + R * q];
          if (apq == 0.0 || fabs(apq) <= kDblEps * sqrt(fabs(app * aqq))) continue;
          const double theta = (aqq - app) / (2.0 * apq);
          double tt;
          if (fabs(theta) > 1e150) tt = 0.5 / theta;
          else {
            tt = 1.0 / (fabs(theta) + sqrt(fma(theta, theta, 1.0)));
            if (theta < 0.0) tt = -tt;
          }
          const double cc = 1.0 / sqrt(fma(tt, tt, 1.0));
          pp[c] = p;
          qq[c] = q;
          cs[c] = cc;
          sn[c] = tt * cc;
          ++c;
        }
        np = c;
        if (c) rotated = 1;
      }
      __syncthreads();
      const int nr = np;
      if (nr == 0) continue;                         // uniform: read after the barrier
      for (int e = t; e < nr * R; e += 256) {        // rows p, q
        const int m = e / R, j = e - m * R;
        const int p = pp[m], q = qq[m];
        const double c = cs[m], s = sn[m];
        const double xx = a[p + R * j], yy = a[q + R * j];
        a[p + R * j] = fma(c, xx, -(s * yy));
        a[q + R * j] = fma(s, xx, c * yy);
      }
      __syncthreads();
      for (int e = t; e < nr * R; e += 256) {        // columns p, q and the vectors
        const int m = e / R, i = e - m * R;
        const int p = pp[m], q = qq[m];
        const double c = cs[m], s = sn[m];
        const double xx = a[i + R * p], yy = a[i + R * q];
        a[i + R * p] = fma(c, xx, -(s * yy));
        a[i + R * q] = fma(s, xx, c * yy);
        const double vx = v[i + R * p], vy = v[i + R * q];
        v[i + R * p] = fma(c, vx, -(s * vy));
        v[i + R * q] = fma(s, vx, c * vy);
      }
      __syncthreads();
      for (int m = t; m < nr; m += 256) {
        a[pp[m] + R * qq[m]] = 0.0;
        a[qq[m] + R * pp[m]] = 0.0;
      }
      __syncthreads();
    }
    __syncthreads();
    if (!rotated) break;                             // uniform
    __syncthreads();
  }
  for (int e = t; e < R * R; e += 256) Q[e] = v[e];
  for (int i = t; i < R; i += 256) evals[i] = a[i + R * i];
}

// T[a][b] = sum_k Q[a][k] sqrt(1 / ev[k]) Q[b][k]  ((M^T M)^{-1/2})
__global__ void k_inv_sqrt(int R, const double *__restrict__ Q, const double *__restrict__ ev, double *__restrict__ T) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= R * R) return;
  const int a = e % R, b = e / R;
  double s = 0.0;
  for (int k = 0; k < R; ++k) s = fma(Q[a + R * k] * sqrt(1.0 / ev[k]), Q[b + R * k], s);
  T[a + R * b] = s;
}

// A[i][c] = sum_k M[i][k] T[k][c] (rows of ldr doubles, columns beyond R left as they are)
__global__ void k_apply_right(int64_t I, int R, int ldr, const double *__restrict__ M, const double *__restrict__ T,
                              double *__restrict__ A) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= I * R) return;
  const int64_t i = e / R;
  const int c = (int)(e - i * R);
  double s = 0.0;
  for (int k = 0; k < R; ++k) s = fma(M[i * ldr + k], T[k + R * c], s);
  A[i * ldr + c] = s;
}

// the Nesterov log fields of a mode updated without NNLS
__global__ void k_params_no_nnls(NnlsParams *p) {
  if (threadIdx.x == 0) {
    p->L0 = 0.0;
    p->mu0 = 0.0;
    p->iters = 0;
    p->fallback = 0;
    p->warm_rows = 0;
    p->warm_steps = 0;
  }
}

}  // namespace

void lu_inverse(hipStream_t s, int R, const double *V, double *Vi) { k_lu_inverse<<<1, 256, 0, s>>>(R, V, Vi); }

void jacobi_eig(hipStream_t s, int R, const double *A, double *evals, double *Q) {
  k_jacobi_eig<<<1, 256, 0, s>>>(R, A, evals, Q);
}

void inv_sqrt_matrix(hipStream_t s, int R, const double *Q, const double *ev, double *T) {
  k_inv_sqrt<<<(R * R + 255) / 256, 256, 0, s>>>(R, Q, ev, T);
}

void apply_right(hipStream_t s, int64_t I, int R, int ldr, const double *M, const double *T, double *A) {
  k_apply_right<<<(unsigned)((I * R + 255) / 256), 256, 0, s>>>(I, R, ldr, M, T, A);
}

void params_no_nnls(hipStream_t s, NnlsParams *p) { k_params_no_nnls<<<1, 64, 0, s>>>(p); }

}  // namespace ptn
