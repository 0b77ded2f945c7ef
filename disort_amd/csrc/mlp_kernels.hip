// Flux surrogate MLP on gfx950: BF16 weights and activations, FP32
// accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16), with
// hand-written forward, backward and Adam kernels.  No counterpart in the
// reference: the network learns DISORT fluxes from columns the batched
// solver generates, and plain DISORT stays the fallback (disort_amd/surrogate.py).
//
// Network: widths w_0 .. w_L (multiples of 64, the caller pads), hidden
// layers  z = h W + b, h' = GELU(z)  (tanh form), last layer linear.
// Parameters per layer: W [w_l x w_{l+1}] row-major FP32 master copy and a
// BF16 copy used by the GEMMs, b [w_{l+1}] FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/disort_amd.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

enum Epi { EPI_GELU = 0, EPI_LINEAR = 1, EPI_WGRAD = 2, EPI_DGELU = 3 };

// C[M x N] = A B^T with A [M][K] and B [N][K] both K-contiguous BF16 (row
// strides lda, ldb); K a multiple of 64, M and N any multiple of 64 (tiles
// are masked).  Every product of the network has this form once the
// weights are kept in both orientations and the training pass keeps the
// transposed activations (see disort_mlp below).
struct GemmArgs {
  const __bf16* A;
  int64_t lda;
  const __bf16* B;
  int64_t ldb;
  int M, N, K;
  int kc;             // K per split (EPI_WGRAD: blockIdx.z selects the split)
  const float* bias;  // EPI_GELU / EPI_LINEAR
  __bf16* Z;          // EPI_GELU: pre-activation (training), may be null
  __bf16* H;          // EPI_GELU: activation; EPI_DGELU: gradient output
  __bf16* HT;         // transposed copy of H [N][ldt] (training), may be null
  int64_t ldt;
  float* F;           // EPI_LINEAR output; EPI_WGRAD: partial sums [split][M][N]
  const __bf16* Zin;  // EPI_DGELU: pre-activation of the layer below
};

// tanh(u) = 1 - 2 / (1 + e^{2u}) with the hardware exponential and a fast
// reciprocal: a few VALU operations instead of the library tanhf, accurate
// to a few FP32 ulp -- far below the BF16 rounding of the stored activation
// (the epilogue, not the matrix cores, bounded the GEMM with tanhf).
__device__ __forceinline__ float fast_tanh(float u) {
  const float e = __expf(2.0f * u);
  return 1.0f - __fdividef(2.0f, 1.0f + e);
}
__device__ __forceinline__ float gelu(float z) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  return 0.5f * z * (1.0f + fast_tanh(k0 * (z + k1 * z * z * z)));
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  const float u = k0 * (z + k1 * z * z * z);
  const float th = fast_tanh(u);
  return 0.5f * (1.0f + th) + 0.5f * z * (1.0f - th * th) * k0 * (1.0f + 3.0f * k1 * z * z);
}

constexpr int BM = 128, BK = 64, LDS_K = BK + 8;  // LDS row pitch: 144 B, conflict-free 16-byte fragment reads

// 128 x BN output tile per 256-thread workgroup, 4 wavefronts in 2 x 2, each
// a 64 x BN/2 block of 4 x BN/32 tiles of v_mfma_f32_16x16x32_bf16 (FP32
// accumulators in registers; BN = 128 needs 156 VGPRs, three workgroups per
// CU).  Per 64-deep K tile the workgroup stages A
// (128 x 64) and B (BN x 64) in LDS as [row][k]; the next tile's global
// loads (16 bytes per lane, coalesced along k) are in flight while the
// current one is multiplied.  MFMA lane layout (16x16x32): lane l supplies
// A[l&15][8(l>>4)..+7] and B[l&15][8(l>>4)..+7]; it holds
// C[4(l>>4)+r][l&15], r = 0..3.
template <int BN, int EPI>
__global__ __launch_bounds__(256, 2) void k_gemm(GemmArgs g) {
  constexpr int WN = BN / 2, TI = 4, TJ = WN / 16;
  constexpr int AL = BM * BK / 8 / 256, BL = BN * BK / 8 / 256;  // 16-byte loads per thread
  __shared__ __attribute__((aligned(16))) __bf16 sA[BM * LDS_K];
  __shared__ __attribute__((aligned(16))) __bf16 sB[BN * LDS_K];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w >> 1, wn = w & 1;
  // XCD-aware order: workgroup b runs on XCD b % 8; give each XCD whole
  // rows of N tiles so one M tile's A rows are reused from that XCD's L2.
  int tm = blockIdx.y, tn = blockIdx.x;
  {
    const int NT = gridDim.x, MT = gridDim.y;
    const int b = blockIdx.y * NT + blockIdx.x;
    if (MT % 8 == 0) {
      const int x = b & 7, loc = b >> 3;
      tn = loc % NT;
      tm = (loc / NT) * 8 + x;
    }
  }
  const int m0 = tm * BM, n0 = tn * BN;
  const int kb = (EPI == EPI_WGRAD) ? blockIdx.z * g.kc : 0;
  const int ke = (EPI == EPI_WGRAD) ? kb + g.kc : g.K;

  // Global source rows of this thread's loads (clamped into the matrix; the
  // rows beyond M or N only feed outputs that are never stored).
  const int lr = t >> 3, lk = (t & 7) * 8;
  const __bf16* ga[AL];
  const __bf16* gb[BL];
#pragma unroll
  for (int i = 0; i < AL; ++i) ga[i] = g.A + (int64_t)min(m0 + lr + 32 * i, g.M - 1) * g.lda + lk;
#pragma unroll
  for (int i = 0; i < BL; ++i) gb[i] = g.B + (int64_t)min(n0 + lr + 32 * i, g.N - 1) * g.ldb + lk;
  bf16x8 ra[AL], rb[BL];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < AL; ++i) ra[i] = *(const bf16x8*)(ga[i] + k0);
#pragma unroll
    for (int i = 0; i < BL; ++i) rb[i] = *(const bf16x8*)(gb[i] + k0);
  };
  auto stage = [&]() {
#pragma unroll
    for (int i = 0; i < AL; ++i) *(bf16x8*)&sA[(lr + 32 * i) * LDS_K + lk] = ra[i];
#pragma unroll
    for (int i = 0; i < BL; ++i) *(bf16x8*)&sB[(lr + 32 * i) * LDS_K + lk] = rb[i];
  };

  f32x4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  load(kb);
  stage();
  __syncthreads();
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  for (int k0 = kb; k0 < ke; k0 += BK) {
    const bool more = k0 + BK < ke;
    if (more) load(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 32) {
      bf16x8 af[TI], bf[TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i) af[i] = *(const bf16x8*)&sA[(wm * 64 + i * 16 + fr) * LDS_K + kk + fk];
#pragma unroll
      for (int j = 0; j < TJ; ++j) bf[j] = *(const bf16x8*)&sB[(wn * WN + j * 16 + fr) * LDS_K + kk + fk];
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
    if (more) {
      stage();
      __syncthreads();
    }
  }

#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      const int row0 = m0 + wm * 64 + i * 16 + (lane >> 4) * 4;
      const int col = n0 + wn * WN + j * 16 + (lane & 15);
      if (row0 >= g.M || col >= g.N) continue;  // M and N are multiples of 16: whole fragments
      bf16x4 ht;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t o = (int64_t)(row0 + r) * g.N + col;
        const float v = acc[i][j][r];
        if (EPI == EPI_GELU) {
          const float z = v + g.bias[col];
          if (g.Z) g.Z[o] = (__bf16)z;
          const __bf16 h = (__bf16)gelu(z);
          g.H[o] = h;
          ht[r] = h;
        } else if (EPI == EPI_LINEAR) {
          g.F[o] = v + g.bias[col];
        } else if (EPI == EPI_WGRAD) {
          g.F[(int64_t)blockIdx.z * g.M * g.N + o] = v;
        } else {
          const __bf16 d = (__bf16)(v * gelu_grad((float)g.Zin[o]));
          g.H[o] = d;
          ht[r] = d;
        }
      }
      if ((EPI == EPI_GELU || EPI == EPI_DGELU) && g.HT) *(bf16x4*)&g.HT[(int64_t)col * g.ldt + row0] = ht;
    }
}

// out[e] = sum_s part[s][e] (the split-K partial sums of a weight gradient).
__global__ void k_sum_splits(const float* part, int S, int64_t n, float* out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float s = 0.f;
  for (int q = 0; q < S; ++q) s += part[q * n + e];
  out[e] = s;
}

// dst [C][R] = src [R][C]^T for BF16, R and C multiples of 64 (64 x 64 tiles through LDS).
__global__ __launch_bounds__(256) void k_transpose(const __bf16* src, int R, int Cn, __bf16* dst) {
  __shared__ __bf16 tile[64][66];
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64, t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = t + 256 * q, r = e >> 6, c = e & 63;
    tile[r][c] = src[(int64_t)(r0 + r) * Cn + c0 + c];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = t + 256 * q, c = e >> 6, r = e & 63;
    dst[(int64_t)(c0 + c) * R + r0 + r] = tile[r][c];
  }
}

// db[n] = sum_b dz^T[n][b]: one 256-thread workgroup per row of the
// transposed gradient (coalesced reads, block reduction).
__global__ __launch_bounds__(256) void k_rowsum(const __bf16* dzt, int B, int N, float* db) {
  __shared__ float part[4];
  const int n = blockIdx.x, t = threadIdx.x;
  float s = 0.f;
  for (int b = t; b < B; b += 256) s += (float)dzt[(int64_t)n * B + b];
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
  if ((t & 63) == 0) part[t >> 6] = s;
  __syncthreads();
  if (t == 0) db[n] = (part[0] + part[1]) + (part[2] + part[3]);
}

// W^T in BF16 ([N][K]) from the FP32 master W ([K][N]).
__global__ void k_weight_t(const float* w, int K, int N, __bf16* wt) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)K * N) return;
  const int64_t k = e / N, n = e % N;
  wt[n * K + k] = (__bf16)w[e];
}

// Loss 1/(B*n_valid) sum mask (y - t)^2 and its gradient dY (BF16): each
// 256-thread workgroup reduces a strip of 16 elements per thread and adds
// one partial sum to *loss.
__global__ __launch_bounds__(256) void k_mse(const float* y, const float* tgt, const float* mask, int B, int N, float scale,
                                             __bf16* dy, float* loss) {
  __shared__ float part[4];
  const int64_t total = (int64_t)B * N;
  const int64_t base = (int64_t)blockIdx.x * 256 * 16 + threadIdx.x;
  float l = 0.f;
#pragma unroll 4
  for (int q = 0; q < 16; ++q) {
    const int64_t e = base + (int64_t)q * 256;
    if (e < total) {
      const int n = (int)(e % N);
      const float d = (y[e] - tgt[e]) * mask[n];
      dy[e] = (__bf16)(2.0f * scale * d);
      l += scale * d * d;
    }
  }
  for (int off = 32; off >= 1; off >>= 1) l += __shfl_xor(l, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, (part[0] + part[1]) + (part[2] + part[3]));
}

// Adam on a contiguous block of parameters; optionally refreshes the BF16
// copy used by the GEMMs.
__global__ void k_adam(float* p, const float* g, float* m, float* v, __bf16* pb, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gr = g[e];
  const float mm = b1 * m[e] + (1.f - b1) * gr;
  const float vv = b2 * v[e] + (1.f - b2) * gr * gr;
  m[e] = mm;
  v[e] = vv;
  const float x = p[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  p[e] = x;
  if (pb) pb[e] = (__bf16)x;
}

__global__ void k_to_bf16(const float* x, __bf16* y, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) y[e] = (__bf16)x[e];
}

unsigned blocks(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct disort_mlp {
  int L = 0, max_batch = 0;
  std::vector<int> w;
  std::vector<int64_t> woff, boff;  // offsets of W_l, b_l in the packed parameter vector
  int64_t nparam = 0;
  float *p = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;  // packed, FP32
  __bf16* wb = nullptr;                                             // BF16 W_l [w_l][w_{l+1}] (same offsets)
  __bf16* wt = nullptr;                                             // BF16 W_l^T [w_{l+1}][w_l] (same offsets)
  std::vector<__bf16*> Z, H, HT;  // per hidden layer: [B x w_{l+1}], H^T [w_{l+1} x B]
  __bf16 *xt = nullptr;           // input of the last forward, transposed [w_0 x B]
  __bf16 *dz0 = nullptr, *dz1 = nullptr, *dzt0 = nullptr, *dzt1 = nullptr, *dy = nullptr, *dyt = nullptr;
  float* part = nullptr;  // split-K partial sums of a weight gradient
  int64_t part_elems = 0;
  float* y = nullptr;
  float* loss = nullptr;
  const __bf16* x = nullptr;  // input of the last forward
  int last_batch = 0;
  std::string err;
};

namespace {

int mfail(disort_mlp* mlp, int code, const std::string& msg) {
  if (mlp) mlp->err = msg;
  return code;
}

#define MCHK(mlp, call)                                                  \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      (mlp)->err = std::string(#call) + ": " + hipGetErrorString(e_);    \
      return DISORT_E_HIP;                                               \
    }                                                                    \
  } while (0)

hipStream_t stream_of(disort_ctx* ctx) { return (hipStream_t)disort_ctx_stream(ctx); }

template <int EPI>
int gemm(disort_mlp* mlp, hipStream_t s, const GemmArgs& g, int splits = 1) {
  if (g.M % 64 || g.N % 64 || g.K % BK || g.M <= 0 || g.N <= 0 || g.K <= 0)
    return mfail(mlp, DISORT_E_INVALID, "GEMM dimensions must be positive multiples of 64");
  // 128 x 128 tiles, three workgroups per CU: the 128 x 256 form (one
  // wavefront per SIMD) measured 0.6x as fast at 65,536 rows.
  const unsigned gm = (unsigned)((g.M + BM - 1) / BM);
  hipLaunchKernelGGL((k_gemm<128, EPI>), dim3((g.N + 127) / 128, gm, splits), dim3(256), 0, s, g);
  MCHK(mlp, hipGetLastError());
  return DISORT_OK;
}

// Refresh the BF16 copies of every weight matrix from the FP32 master.
int refresh_bf16(disort_mlp* mlp, hipStream_t s) {
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(mlp->nparam)), dim3(256), 0, s, mlp->p, mlp->wb, mlp->nparam);
  for (int l = 0; l < mlp->L; ++l)
    hipLaunchKernelGGL(k_weight_t, dim3(blocks((int64_t)mlp->w[l] * mlp->w[l + 1])), dim3(256), 0, s,
                       mlp->p + mlp->woff[l], mlp->w[l], mlp->w[l + 1], mlp->wt + mlp->woff[l]);
  MCHK(mlp, hipGetLastError());
  return DISORT_OK;
}

// Forward through every layer; `train` keeps Z, H^T and x^T for a backward.
int forward(disort_ctx* ctx, disort_mlp* mlp, const uint16_t* x, int B, float* y, bool train) {
  if (!ctx || !mlp || !x || !y || B <= 0 || B % 64 || B > mlp->max_batch)
    return mfail(mlp, DISORT_E_INVALID, "batch must be a positive multiple of 64 within max_batch");
  hipStream_t s = stream_of(ctx);
  const __bf16* h = (const __bf16*)x;
  int rc;
  if (train) {
    hipLaunchKernelGGL(k_transpose, dim3(mlp->w[0] / 64, B / 64), dim3(256), 0, s, h, B, mlp->w[0], mlp->xt);
    MCHK(mlp, hipGetLastError());
  }
  for (int l = 0; l < mlp->L; ++l) {
    GemmArgs g{};
    g.A = h;
    g.lda = mlp->w[l];
    g.B = mlp->wt + mlp->woff[l];  // W_l^T: [w_{l+1}][w_l]
    g.ldb = mlp->w[l];
    g.M = B;
    g.N = mlp->w[l + 1];
    g.K = mlp->w[l];
    g.bias = mlp->p + mlp->boff[l];
    if (l + 1 < mlp->L) {
      g.Z = train ? mlp->Z[l] : nullptr;
      g.H = mlp->H[l];
      g.HT = train ? mlp->HT[l] : nullptr;
      g.ldt = B;
      if ((rc = gemm<EPI_GELU>(mlp, s, g))) return rc;
      h = mlp->H[l];
    } else {
      g.F = y;
      if ((rc = gemm<EPI_LINEAR>(mlp, s, g))) return rc;
    }
  }
  if (train) {
    mlp->x = (const __bf16*)x;
    mlp->last_batch = B;
  }
  return DISORT_OK;
}

// Backward given dz of the output layer and its transpose dzt [w_L x B].
int backward(disort_ctx* ctx, disort_mlp* mlp, const __bf16* dz, const __bf16* dzt) {
  hipStream_t s = stream_of(ctx);
  const int B = mlp->last_batch;
  int rc;
  for (int l = mlp->L - 1; l >= 0; --l) {
    const int K = mlp->w[l], N = mlp->w[l + 1];
    // dW_l = h_l^T dz = (H^T_l) (dz^T)^T over the batch, split along the batch.
    GemmArgs gw{};
    gw.A = (l == 0) ? mlp->xt : mlp->HT[l - 1];
    gw.lda = B;
    gw.B = dzt;
    gw.ldb = B;
    gw.M = K;
    gw.N = N;
    gw.K = B;
    const int tiles = ((K + BM - 1) / BM) * ((N + 127) / 128);
    int splits = 1;
    while (splits * 2 <= B / BK && (B / BK) % (splits * 2) == 0 && tiles * splits < 256 &&
           (int64_t)splits * 2 * K * N <= mlp->part_elems)
      splits *= 2;
    gw.kc = B / splits;
    gw.F = mlp->part;
    if ((rc = gemm<EPI_WGRAD>(mlp, s, gw, splits))) return rc;
    hipLaunchKernelGGL(k_sum_splits, dim3(blocks((int64_t)K * N)), dim3(256), 0, s, mlp->part, splits,
                       (int64_t)K * N, mlp->grad + mlp->woff[l]);
    hipLaunchKernelGGL(k_rowsum, dim3(N), dim3(256), 0, s, dzt, B, N, mlp->grad + mlp->boff[l]);
    MCHK(mlp, hipGetLastError());
    if (l > 0) {  // dz_{l-1} = (dz W_l^T) * GELU'(z_{l-1}), and its transpose
      __bf16* next = (dz == mlp->dz0) ? mlp->dz1 : mlp->dz0;
      __bf16* nextt = (dz == mlp->dz0) ? mlp->dzt1 : mlp->dzt0;
      GemmArgs gd{};
      gd.A = dz;
      gd.lda = N;
      gd.B = mlp->wb + mlp->woff[l];  // W_l: [w_l][w_{l+1}]
      gd.ldb = N;
      gd.M = B;
      gd.N = K;
      gd.K = N;
      gd.H = next;
      gd.HT = nextt;
      gd.ldt = B;
      gd.Zin = mlp->Z[l - 1];
      if ((rc = gemm<EPI_DGELU>(mlp, s, gd))) return rc;
      dz = next;
      dzt = nextt;
    }
  }
  return DISORT_OK;
}

}  // namespace

extern "C" {

int disort_mlp_create(disort_ctx* ctx, int n_layers, const int* widths, int max_batch, disort_mlp** out) {
  if (!ctx || !widths || !out || n_layers < 1 || max_batch < 64 || max_batch % 64) return DISORT_E_INVALID;
  for (int l = 0; l <= n_layers; ++l)
    if (widths[l] <= 0 || widths[l] % 64) return DISORT_E_INVALID;
  if (hipSetDevice(disort_ctx_device(ctx)) != hipSuccess) return DISORT_E_HIP;
  auto* mlp = new disort_mlp();
  mlp->L = n_layers;
  mlp->max_batch = max_batch;
  mlp->w.assign(widths, widths + n_layers + 1);
  int64_t off = 0, wprod = 0;
  for (int l = 0; l < n_layers; ++l) {
    mlp->woff.push_back(off);
    off += (int64_t)widths[l] * widths[l + 1];
    wprod = std::max<int64_t>(wprod, (int64_t)widths[l] * widths[l + 1]);
    mlp->boff.push_back(off);
    off += widths[l + 1];
  }
  mlp->nparam = off;
  mlp->part_elems = 64 * wprod;
  auto alloc = [&](void** p, size_t bytes) { return hipMalloc(p, bytes) == hipSuccess; };
  bool ok = alloc((void**)&mlp->p, off * 4) && alloc((void**)&mlp->grad, off * 4) && alloc((void**)&mlp->m, off * 4) &&
            alloc((void**)&mlp->v, off * 4) && alloc((void**)&mlp->wb, off * 2) && alloc((void**)&mlp->wt, off * 2) &&
            alloc((void**)&mlp->part, (size_t)mlp->part_elems * 4);
  int wmax = 0;
  for (int x : mlp->w) wmax = wmax > x ? wmax : x;
  for (int l = 0; l + 1 < n_layers && ok; ++l) {
    __bf16 *z = nullptr, *h = nullptr, *ht = nullptr;
    const size_t e = (size_t)max_batch * widths[l + 1] * 2;
    ok = alloc((void**)&z, e) && alloc((void**)&h, e) && alloc((void**)&ht, e);
    mlp->Z.push_back(z);
    mlp->H.push_back(h);
    mlp->HT.push_back(ht);
  }
  const size_t ew = (size_t)max_batch * wmax * 2;
  ok = ok && alloc((void**)&mlp->xt, (size_t)max_batch * widths[0] * 2) && alloc((void**)&mlp->dz0, ew) &&
       alloc((void**)&mlp->dz1, ew) && alloc((void**)&mlp->dzt0, ew) && alloc((void**)&mlp->dzt1, ew) &&
       alloc((void**)&mlp->dy, (size_t)max_batch * widths[n_layers] * 2) &&
       alloc((void**)&mlp->dyt, (size_t)max_batch * widths[n_layers] * 2) &&
       alloc((void**)&mlp->y, (size_t)max_batch * widths[n_layers] * 4) && alloc((void**)&mlp->loss, 4);
  if (!ok) {
    disort_mlp_destroy(ctx, mlp);
    return DISORT_E_NOMEM;
  }
  (void)hipMemset(mlp->m, 0, off * 4);
  (void)hipMemset(mlp->v, 0, off * 4);
  (void)hipMemset(mlp->p, 0, off * 4);
  (void)hipMemset(mlp->wb, 0, off * 2);
  (void)hipMemset(mlp->wt, 0, off * 2);
  *out = mlp;
  return DISORT_OK;
}

int disort_mlp_destroy(disort_ctx* ctx, disort_mlp* mlp) {
  if (!mlp) return DISORT_OK;
  if (ctx) (void)hipStreamSynchronize(stream_of(ctx));
  for (void* q : {(void*)mlp->p, (void*)mlp->grad, (void*)mlp->m, (void*)mlp->v, (void*)mlp->wb, (void*)mlp->wt,
                  (void*)mlp->part, (void*)mlp->xt, (void*)mlp->dz0, (void*)mlp->dz1, (void*)mlp->dzt0,
                  (void*)mlp->dzt1, (void*)mlp->dy, (void*)mlp->dyt, (void*)mlp->y, (void*)mlp->loss})
    if (q) (void)hipFree(q);
  for (auto* q : mlp->Z) if (q) (void)hipFree(q);
  for (auto* q : mlp->H) if (q) (void)hipFree(q);
  for (auto* q : mlp->HT) if (q) (void)hipFree(q);
  delete mlp;
  return DISORT_OK;
}

const char* disort_mlp_last_error(const disort_mlp* mlp) { return mlp ? mlp->err.c_str() : "null mlp"; }

int64_t disort_mlp_param_count(const disort_mlp* mlp) { return mlp ? mlp->nparam : 0; }

int disort_mlp_set_params(disort_ctx* ctx, disort_mlp* mlp, const float* host) {
  if (!ctx || !mlp || !host) return DISORT_E_INVALID;
  hipStream_t s = stream_of(ctx);
  MCHK(mlp, hipMemcpyAsync(mlp->p, host, mlp->nparam * 4, hipMemcpyHostToDevice, s));
  int rc;
  if ((rc = refresh_bf16(mlp, s))) return rc;
  MCHK(mlp, hipMemsetAsync(mlp->m, 0, mlp->nparam * 4, s));
  MCHK(mlp, hipMemsetAsync(mlp->v, 0, mlp->nparam * 4, s));
  MCHK(mlp, hipStreamSynchronize(s));
  return DISORT_OK;
}

int disort_mlp_get_params(disort_ctx* ctx, disort_mlp* mlp, float* host) {
  if (!ctx || !mlp || !host) return DISORT_E_INVALID;
  MCHK(mlp, hipMemcpyAsync(host, mlp->p, mlp->nparam * 4, hipMemcpyDeviceToHost, stream_of(ctx)));
  MCHK(mlp, hipStreamSynchronize(stream_of(ctx)));
  return DISORT_OK;
}

int disort_mlp_get_grads(disort_ctx* ctx, disort_mlp* mlp, float* host) {
  if (!ctx || !mlp || !host) return DISORT_E_INVALID;
  MCHK(mlp, hipMemcpyAsync(host, mlp->grad, mlp->nparam * 4, hipMemcpyDeviceToHost, stream_of(ctx)));
  MCHK(mlp, hipStreamSynchronize(stream_of(ctx)));
  return DISORT_OK;
}

// y[B x w_L] (FP32, device) = network(x[B x w_0] BF16 bits, device); keeps
// what a following disort_mlp_backward needs.
int disort_mlp_forward(disort_ctx* ctx, disort_mlp* mlp, const uint16_t* x, int B, float* y) {
  return forward(ctx, mlp, x, B, y, true);
}

// The same output without keeping anything for a backward pass (serving).
int disort_mlp_infer(disort_ctx* ctx, disort_mlp* mlp, const uint16_t* x, int B, float* y) {
  return forward(ctx, mlp, x, B, y, false);
}

// Gradients of the parameters for output gradient dy [B x w_L] (BF16 bits,
// device) of the last forward, into the packed FP32 gradient vector.
int disort_mlp_backward(disort_ctx* ctx, disort_mlp* mlp, const uint16_t* dy) {
  if (!ctx || !mlp || !dy || !mlp->x) return mfail(mlp, DISORT_E_INVALID, "backward needs a preceding forward");
  hipStream_t s = stream_of(ctx);
  const int B = mlp->last_batch, wL = mlp->w[mlp->L];
  hipLaunchKernelGGL(k_transpose, dim3(wL / 64, B / 64), dim3(256), 0, s, (const __bf16*)dy, B, wL, mlp->dyt);
  MCHK(mlp, hipGetLastError());
  return backward(ctx, mlp, (const __bf16*)dy, mlp->dyt);
}

// One Adam step (t >= 1) on the current gradients.
int disort_mlp_adam(disort_ctx* ctx, disort_mlp* mlp, float lr, float beta1, float beta2, float eps, int t) {
  if (!ctx || !mlp || t < 1) return DISORT_E_INVALID;
  hipStream_t s = stream_of(ctx);
  const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3(blocks(mlp->nparam)), dim3(256), 0, s, mlp->p, mlp->grad, mlp->m, mlp->v, mlp->wb,
                     mlp->nparam, lr, beta1, beta2, eps, c1, c2);
  for (int l = 0; l < mlp->L; ++l)
    hipLaunchKernelGGL(k_weight_t, dim3(blocks((int64_t)mlp->w[l] * mlp->w[l + 1])), dim3(256), 0, s,
                       mlp->p + mlp->woff[l], mlp->w[l], mlp->w[l + 1], mlp->wt + mlp->woff[l]);
  MCHK(mlp, hipGetLastError());
  return DISORT_OK;
}

// forward + masked mean-squared error against tgt [B x w_L] (FP32, device)
// + backward + Adam.  The loss of the batch is accumulated into the
// device float *loss_dev (the caller clears it).
int disort_mlp_train_step(disort_ctx* ctx, disort_mlp* mlp, const uint16_t* x, const float* tgt, const float* mask_dev,
                          int B, int n_valid, float lr, float beta1, float beta2, float eps, int t, float* loss_dev) {
  int rc;
  if ((rc = disort_mlp_forward(ctx, mlp, x, B, mlp->y))) return rc;
  hipStream_t s = stream_of(ctx);
  const int N = mlp->w[mlp->L];
  const float scale = 1.0f / ((float)B * (float)(n_valid > 0 ? n_valid : N));
  hipLaunchKernelGGL(k_mse, dim3(blocks((int64_t)B * N, 256 * 16)), dim3(256), 0, s, mlp->y, tgt, mask_dev, B, N, scale,
                     mlp->dy, loss_dev);
  MCHK(mlp, hipGetLastError());
  if ((rc = disort_mlp_backward(ctx, mlp, (const uint16_t*)mlp->dy))) return rc;
  return disort_mlp_adam(ctx, mlp, lr, beta1, beta2, eps, t);
}

}  // extern "C"
