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

// Value of x in lane `lane` (wave-uniform) of the calling wavefront.
__device__ __forceinline__ double lane_bcast(double x, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), lane),
                          __builtin_amdgcn_readlane(__double2loint(x), lane));
}

// Back substitution over the M panels of one boundary problem, the whole
// wavefront, lane k = U row k of each panel (N <= 64; other lanes repeat row
// 0 and store nothing):
//   x_k = (r_k - sum_c V(k,c) xnext_c - sum_{c>k} U(k,c) x_c) / U(k,k),
// with the panel stored as [U with 1/U(k,k) on its diagonal | V | r], U row
// k either in place or, with Shifted, starting at its diagonal (entry (k, c)
// at column c - k).  The x_c of other lanes come by readlane at a
// compile-time lane index, so the column chain waits on no memory access and
// every load of the panel can be issued up front.
template <int N, bool Shifted = false>
__device__ __forceinline__ void panel_back_substitute(const double* panel, int64_t t0, int M, double* cvec) {
  static_assert(N <= 64, "one U row per lane");
  constexpr int PS = panel_stride(N);
  const int lane = threadIdx.x & 63;
  const int k = lane < N ? lane : 0;
  wave_lds_order();  // the panel rows were stored by other lanes of this wavefront
  double xn = 0.0;   // this lane's entry of the layer below
  for (int ip = M - 1; ip >= 0; --ip) {
    const int64_t tt = t0 + ip;
    const double* pp = panel + tt * PS;
    double v = pp[2 * N * N + k];
    if (ip < M - 1) {
      static_for<0, N>([&](auto cj) {
        constexpr int c = decltype(cj)::value;
        v -= pp[N * N + k * N + c] * lane_bcast(xn, c);
      });
    }
    const double rdg = Shifted ? pp[k * N] : pp[k * N + k];
    double xk = 0.0;
    static_for<0, N>([&](auto cj) {
      constexpr int c = N - 1 - decltype(cj)::value;
      const double xc = lane_bcast(v * rdg, c);
      const double u = Shifted ? pp[k * N + (c > k ? c - k : 0)] : pp[k * N + c];
      xk = (k == c) ? xc : xk;
      v = (k < c) ? v - u * xc : v;
    });
    xn = xk;
    if (lane < N) cvec[tt * N + k] = xk;
  }
}

// ---------------------------------------------------------------------------
// Block-cooperative vector kernels on LDS (all threads call them).
// ---------------------------------------------------------------------------
// With OW (a one-wavefront block) the serial chains run in registers: the
// solved entry reaches the other lanes by readlane and the barriers become
// wave-local ordering points; the operations and their order are those of
// the block form.

// x <- M^{-1} x, M lower triangular (pitch ld).  Entry r accumulates its
// subtractions in ascending column order, then divides (as a row-oriented
// forward substitution would).
template <int NN, bool OW = false>
__device__ void blk_lsolve(const double* M, int ld, double* x) {
  const int tid = threadIdx.x;
  double xr = (tid < NN) ? x[tid] : 0.0;
  if constexpr (OW) {
    const int tr = tid < NN ? tid : 0;
    static_for<0, NN>([&](auto ri) {
      constexpr int r = decltype(ri)::value;
      const double xs = lane_bcast(xr, r) / M[r * ld + r];
      const double mr = M[tr * ld + r];
      if (tid == r) xr = xs;
      if (tid > r && tid < NN) xr -= mr * xs;
    });
    wave_lds_order();
    if (tid < NN) x[tid] = xr;
    wave_lds_order();
  } else {
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
}

// x <- M^{-T} x, M lower triangular.
template <int NN, bool OW = false>
__device__ void blk_ltsolve(const double* M, int ld, double* x) {
  const int tid = threadIdx.x;
  double xr = (tid < NN) ? x[tid] : 0.0;
  if constexpr (OW) {
    const int tr = tid < NN ? tid : 0;
    static_for<0, NN>([&](auto ri) {
      constexpr int r = NN - 1 - decltype(ri)::value;
      const double xs = lane_bcast(xr, r) / M[r * ld + r];
      const double mr = M[r * ld + tr];
      if (tid == r) xr = xs;
      if (tid < r) xr -= mr * xs;
    });
    wave_lds_order();
    if (tid < NN) x[tid] = xr;
    wave_lds_order();
  } else {
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
}

// U = M^{-T} V column by column, M lower triangular (pitch ld), thread j < NN
// holding column j in registers: U(r,j) = (V(r,j) - sum_{m>r} M(m,r) U(m,j))
// / M(r,r), the subtractions in ascending m.
template <int NN>
__device__ void blk_ltsolve_cols(const double* M, int ld, const double* V, double* U) {
  const int j = threadIdx.x;
  if (j < NN) {
    double u[NN];
    static_for<0, NN>([&](auto ri) {
      constexpr int r = NN - 1 - decltype(ri)::value;
      double s = V[r * ld + j];
      static_for<r + 1, NN>([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        s -= M[m * ld + r] * u[m];
      });
      u[r] = s / M[r * ld + r];
    });
#pragma unroll
    for (int r = 0; r < NN; ++r) U[r * ld + j] = u[r];
  }
}

// Ordering point of the block: a barrier, or with one wavefront only the
// compiler's (its LDS accesses execute in program order).
template <bool OW>
__device__ __forceinline__ void blk_sync() {
  if constexpr (OW) wave_lds_order();
  else __syncthreads();
}

// x <- M x (lower) or M^T x, through scratch y.
template <int NN, bool TRANS, bool OW = false>
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
  blk_sync<OW>();
  if (tid < NN) x[tid] = y[tid];
  blk_sync<OW>();
}

// x <- V diag(d) V^T x (V column eigenvectors, pitch ld), scratch y.
template <int NN, bool OW = false>
__device__ void blk_vdv(const double* V, int ld, const double* d, double* x, double* y) {
  const int tid = threadIdx.x;
  if (tid < NN) {
    double s = 0.0;
    for (int r = 0; r < NN; ++r) s += V[r * ld + tid] * x[r];
    y[tid] = s * d[tid];
  }
  blk_sync<OW>();
  if (tid < NN) {
    double s = 0.0;
    for (int j = 0; j < NN; ++j) s += V[tid * ld + j] * y[j];
    x[tid] = s;
  }
  blk_sync<OW>();
}

// In-place Cholesky of the lower triangle of M (left-looking, column by
// column, the operation order of the register kernel).  Returns 0 when a
// pivot is not positive (the factor then uses 1 in its place).  With one
// wavefront the lane of row c forms the diagonal in the same pass as the
// rows below it.
template <int NN, bool OW = false>
__device__ int blk_cholesky(double* M, int ld, int* flag) {
  const int tid = threadIdx.x;
  if constexpr (OW) {
    bool ok = true;
    for (int c = 0; c < NN; ++c) {
      const int r = tid >= c && tid < NN ? tid : c;
      double s = M[r * ld + c];
      for (int k = 0; k < c; ++k) s -= M[r * ld + k] * M[c * ld + k];
      const double x = lane_bcast(s, c);
      ok = ok && (x > 0.0);
      const double lcc = sqrt(x > 0.0 ? x : 1.0);
      wave_lds_order();
      if (tid == c) M[c * ld + c] = lcc;
      else if (tid > c && tid < NN) M[r * ld + c] = s / lcc;
      wave_lds_order();
    }
    return ok ? 1 : 0;
  } else {
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
}

// Block-wide sum of one double per thread (every lane of a wavefront ends
// with the same butterfly sum; with one wavefront that is the result).
template <bool OW = false>
__device__ double blk_sum(double v, double* red) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  if constexpr (OW) return v;
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
