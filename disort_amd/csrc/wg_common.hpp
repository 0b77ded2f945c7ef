// Workgroup building blocks shared by the flux workgroup kernels
// (flux_wide.hip) and the general-solver kernels (general_kernels.hip):
// launch geometry per half stream count and block-cooperative LDS vector
// kernels (triangular solves, products, Cholesky, block sums).
#pragma once
#include "flux_common.hpp"

namespace dsa {
namespace {

template <int NN>
struct WideCfg {
  static constexpr int N = 2 * NN;
  static constexpr int LT = NN >= 32 ? 256 : 64;  // k_layer_wg threads: one wavefront (wave-local barriers) up to N = 32, four at N = 64
  static constexpr int LW = NN >= 32 ? 5 : (NN >= 16 ? 3 : 1);  // k_layer_wg waves per SIMD asked for
  static constexpr int R = 3 * NN;                // panel rows
  static constexpr int RP = (R + 63) / 64 * 64;   // rows padded to whole waves
  static constexpr int BT = 2 * RP;               // k_boundary_wg threads
  static constexpr int LD = NN + 1;               // LDS row pitch of NN x NN matrices
};

// ---------------------------------------------------------------------------
// Block-cooperative vector kernels on LDS (all threads call them).
// ---------------------------------------------------------------------------
// x <- M^{-1} x, M lower triangular (pitch ld).  Entry r accumulates its
// subtractions in ascending column order, then divides (as a row-oriented
// forward substitution would).
template <int NN>
__device__ void blk_lsolve(const double* M, int ld, double* x) {
  const int tid = threadIdx.x;
  double xr = (tid < NN) ? x[tid] : 0.0;
  for (int r = 0; r < NN; ++r) {
    if (tid == r) {
      xr = xr / M[r * ld + r];
      x[r] = xr;
    }
    __syncthreads();
    if (tid > r && tid < NN) xr -= M[tid * ld + r] * x[r];
  }
  __syncthreads();
}

// x <- M^{-T} x, M lower triangular.
template <int NN>
__device__ void blk_ltsolve(const double* M, int ld, double* x) {
  const int tid = threadIdx.x;
  double xr = (tid < NN) ? x[tid] : 0.0;
  for (int r = NN - 1; r >= 0; --r) {
    if (tid == r) {
      xr = xr / M[r * ld + r];
      x[r] = xr;
    }
    __syncthreads();
    if (tid < r) xr -= M[r * ld + tid] * x[r];
  }
  __syncthreads();
}

// x <- M x (lower) or M^T x, through scratch y.
template <int NN, bool TRANS>
__device__ void blk_lmul(const double* M, int ld, double* x, double* y) {
  const int tid = threadIdx.x;
  if (tid < NN) {
    double s = 0.0;
    if (!TRANS)
      for (int m = 0; m <= tid; ++m) s += M[tid * ld + m] * x[m];
    else
      for (int m = tid; m < NN; ++m) s += M[m * ld + tid] * x[m];
    y[tid] = s;
  }
  __syncthreads();
  if (tid < NN) x[tid] = y[tid];
  __syncthreads();
}

// x <- V diag(d) V^T x (V column eigenvectors, pitch ld), scratch y.
template <int NN>
__device__ void blk_vdv(const double* V, int ld, const double* d, double* x, double* y) {
  const int tid = threadIdx.x;
  if (tid < NN) {
    double s = 0.0;
    for (int r = 0; r < NN; ++r) s += V[r * ld + tid] * x[r];
    y[tid] = s * d[tid];
  }
  __syncthreads();
  if (tid < NN) {
    double s = 0.0;
    for (int j = 0; j < NN; ++j) s += V[tid * ld + j] * y[j];
    x[tid] = s;
  }
  __syncthreads();
}

// In-place Cholesky of the lower triangle of M (left-looking, column by
// column, the operation order of the register kernel).  Returns 0 when a
// pivot is not positive (the factor then uses 1 in its place).
template <int NN>
__device__ int blk_cholesky(double* M, int ld, int* flag) {
  const int tid = threadIdx.x;
  if (tid == 0) *flag = 1;
  __syncthreads();
  for (int c = 0; c < NN; ++c) {
    if (tid == 0) {
      double x = M[c * ld + c];
      for (int k = 0; k < c; ++k) x -= M[c * ld + k] * M[c * ld + k];
      if (!(x > 0.0)) *flag = 0;
      M[c * ld + c] = sqrt(x > 0.0 ? x : 1.0);
    }
    __syncthreads();
    const double lcc = M[c * ld + c];
    for (int r = c + 1 + tid; r < NN; r += blockDim.x) {
      double s = M[r * ld + c];
      for (int k = 0; k < c; ++k) s -= M[r * ld + k] * M[c * ld + k];
      M[r * ld + c] = s / lcc;
    }
    __syncthreads();
  }
  return *flag;
}

// Block-wide sum of one double per thread.
__device__ double blk_sum(double v, double* red) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < nw; ++k) s += red[k];
  return s;
}

}  // namespace
}  // namespace dsa
