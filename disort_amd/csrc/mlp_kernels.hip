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

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/disort_amd.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

enum Epi { EPI_GELU = 0, EPI_LINEAR = 1, EPI_WGRAD = 2, EPI_DGELU = 3 };

// C[M x N] = A[M x K] B[K x N]; A(m,k) at A[m*sam + k*sak], B(k,n) at
// B[k*sbk + n*sbn]; one of each pair of strides is 1.  M, N multiples of
// 64, K a multiple of 32.
struct GemmArgs {
  const __bf16* A;
  int64_t sam, sak;
  const __bf16* B;
  int64_t sbk, sbn;
  int M, N, K;
  const float* bias;  // EPI_GELU / EPI_LINEAR
  __bf16* Z;          // EPI_GELU: pre-activation
  __bf16* H;          // EPI_GELU: activation; EPI_DGELU: gradient output
  float* F;           // EPI_LINEAR / EPI_WGRAD output
  const __bf16* Zin;  // EPI_DGELU: pre-activation of the layer below
};

__device__ __forceinline__ float gelu(float z) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  return 0.5f * z * (1.0f + tanhf(k0 * (z + k1 * z * z * z)));
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  const float u = k0 * (z + k1 * z * z * z);
  const float th = tanhf(u);
  return 0.5f * (1.0f + th) + 0.5f * z * (1.0f - th * th) * k0 * (1.0f + 3.0f * k1 * z * z);
}

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;

// 64 x 64 output tile per 256-thread workgroup; wave w owns the 32 x 32
// quadrant (w / 2, w % 2) as 2 x 2 MFMA 16x16 tiles.  A and B tiles are
// staged through LDS as [m][k] and [n][k] so each lane's fragment (8
// consecutive k) is one 16-byte LDS read.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) __bf16 As[BM][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[BN][LDK];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < g.K; k0 += BK) {
    if (g.sak == 1) {
      const int m = t >> 2, kk = (t & 3) * 8;
      *(bf16x8*)&As[m][kk] = *(const bf16x8*)(g.A + (int64_t)(m0 + m) * g.sam + (k0 + kk));
    } else {
      const int kk = t >> 3, m = (t & 7) * 8;
      const bf16x8 v = *(const bf16x8*)(g.A + (int64_t)(m0 + m) + (int64_t)(k0 + kk) * g.sak);
#pragma unroll
      for (int e = 0; e < 8; ++e) As[m + e][kk] = v[e];
    }
    if (g.sbn == 1) {
      const int kk = t >> 3, n = (t & 7) * 8;
      const bf16x8 v = *(const bf16x8*)(g.B + (int64_t)(k0 + kk) * g.sbk + (n0 + n));
#pragma unroll
      for (int e = 0; e < 8; ++e) Bs[n + e][kk] = v[e];
    } else {
      const int n = t >> 2, kk = (t & 3) * 8;
      *(bf16x8*)&Bs[n][kk] = *(const bf16x8*)(g.B + (int64_t)(k0 + kk) + (int64_t)(n0 + n) * g.sbn);
    }
    __syncthreads();
    const int kq = (lane >> 4) * 8;
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) af[i] = *(const bf16x8*)&As[wm * 32 + i * 16 + (lane & 15)][kq];
#pragma unroll
    for (int j = 0; j < 2; ++j) bfr[j] = *(const bf16x8*)&Bs[wn * 32 + j * 16 + (lane & 15)][kq];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 32 + i * 16 + (lane >> 4) * 4 + r;
        const int col = n0 + wn * 32 + j * 16 + (lane & 15);
        const int64_t o = (int64_t)row * g.N + col;
        const float v = acc[i][j][r];
        if (EPI == EPI_GELU) {
          const float z = v + g.bias[col];
          g.Z[o] = (__bf16)z;
          g.H[o] = (__bf16)gelu(z);
        } else if (EPI == EPI_LINEAR) {
          g.F[o] = v + g.bias[col];
        } else if (EPI == EPI_WGRAD) {
          g.F[o] = v;
        } else {
          g.H[o] = (__bf16)(v * gelu_grad((float)g.Zin[o]));
        }
      }
}

// Loss 1/(B*n_valid) sum mask (y - t)^2 and its gradient dY (BF16).
__global__ void k_mse(const float* y, const float* tgt, const float* mask, int B, int N, float scale, __bf16* dy,
                      float* loss) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (e < (int64_t)B * N) {
    const int n = (int)(e % N);
    const float d = (y[e] - tgt[e]) * mask[n];
    dy[e] = (__bf16)(2.0f * scale * d);
    l = scale * d * d;
  }
  for (int off = 32; off >= 1; off >>= 1) l += __shfl_xor(l, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}

// db[n] = sum_b dZ[b][n].
__global__ void k_colsum(const __bf16* dz, int B, int N, float* db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int b = 0; b < B; ++b) s += (float)dz[(int64_t)b * N + n];
  db[n] = s;
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
  __bf16* wb = nullptr;                                             // BF16 copy (same offsets)
  std::vector<__bf16*> Z, H;                                        // per hidden layer [B x w_{l+1}]
  __bf16 *dz0 = nullptr, *dz1 = nullptr, *dy = nullptr;
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
int gemm(disort_mlp* mlp, hipStream_t s, const GemmArgs& g) {
  if (g.M % BM || g.N % BN || g.K % BK) return mfail(mlp, DISORT_E_INVALID, "GEMM dimensions must be multiples of 64/64/32");
  hipLaunchKernelGGL(k_gemm<EPI>, dim3(g.N / BN, g.M / BM), dim3(256), 0, s, g);
  MCHK(mlp, hipGetLastError());
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
  int64_t off = 0;
  for (int l = 0; l < n_layers; ++l) {
    mlp->woff.push_back(off);
    off += (int64_t)widths[l] * widths[l + 1];
    mlp->boff.push_back(off);
    off += widths[l + 1];
  }
  mlp->nparam = off;
  auto alloc = [&](void** p, size_t bytes) { return hipMalloc(p, bytes) == hipSuccess; };
  bool ok = alloc((void**)&mlp->p, off * 4) && alloc((void**)&mlp->grad, off * 4) && alloc((void**)&mlp->m, off * 4) &&
            alloc((void**)&mlp->v, off * 4) && alloc((void**)&mlp->wb, off * 2);
  int wmax = 0;
  for (int x : mlp->w) wmax = wmax > x ? wmax : x;
  for (int l = 0; l + 1 < n_layers && ok; ++l) {
    __bf16 *z = nullptr, *h = nullptr;
    ok = alloc((void**)&z, (size_t)max_batch * widths[l + 1] * 2) && alloc((void**)&h, (size_t)max_batch * widths[l + 1] * 2);
    mlp->Z.push_back(z);
    mlp->H.push_back(h);
  }
  ok = ok && alloc((void**)&mlp->dz0, (size_t)max_batch * wmax * 2) && alloc((void**)&mlp->dz1, (size_t)max_batch * wmax * 2) &&
       alloc((void**)&mlp->dy, (size_t)max_batch * widths[n_layers] * 2) &&
       alloc((void**)&mlp->y, (size_t)max_batch * widths[n_layers] * 4) && alloc((void**)&mlp->loss, 4);
  if (!ok) {
    disort_mlp_destroy(ctx, mlp);
    return DISORT_E_NOMEM;
  }
  (void)hipMemset(mlp->m, 0, off * 4);
  (void)hipMemset(mlp->v, 0, off * 4);
  (void)hipMemset(mlp->p, 0, off * 4);
  (void)hipMemset(mlp->wb, 0, off * 2);
  *out = mlp;
  return DISORT_OK;
}

int disort_mlp_destroy(disort_ctx* ctx, disort_mlp* mlp) {
  if (!mlp) return DISORT_OK;
  if (ctx) (void)hipStreamSynchronize(stream_of(ctx));
  for (void* q : {(void*)mlp->p, (void*)mlp->grad, (void*)mlp->m, (void*)mlp->v, (void*)mlp->wb, (void*)mlp->dz0,
                  (void*)mlp->dz1, (void*)mlp->dy, (void*)mlp->y, (void*)mlp->loss})
    if (q) (void)hipFree(q);
  for (auto* q : mlp->Z) if (q) (void)hipFree(q);
  for (auto* q : mlp->H) if (q) (void)hipFree(q);
  delete mlp;
  return DISORT_OK;
}

const char* disort_mlp_last_error(const disort_mlp* mlp) { return mlp ? mlp->err.c_str() : "null mlp"; }

int64_t disort_mlp_param_count(const disort_mlp* mlp) { return mlp ? mlp->nparam : 0; }

int disort_mlp_set_params(disort_ctx* ctx, disort_mlp* mlp, const float* host) {
  if (!ctx || !mlp || !host) return DISORT_E_INVALID;
  hipStream_t s = stream_of(ctx);
  MCHK(mlp, hipMemcpyAsync(mlp->p, host, mlp->nparam * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(mlp->nparam)), dim3(256), 0, s, mlp->p, mlp->wb, mlp->nparam);
  MCHK(mlp, hipGetLastError());
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

// y[B x w_L] (FP32, device) = network(x[B x w_0] BF16 bits, device).
int disort_mlp_forward(disort_ctx* ctx, disort_mlp* mlp, const uint16_t* x, int B, float* y) {
  if (!ctx || !mlp || !x || !y || B <= 0 || B % 64 || B > mlp->max_batch)
    return mfail(mlp, DISORT_E_INVALID, "batch must be a positive multiple of 64 within max_batch");
  hipStream_t s = stream_of(ctx);
  const __bf16* h = (const __bf16*)x;
  int rc;
  for (int l = 0; l < mlp->L; ++l) {
    GemmArgs g{};
    g.A = h;
    g.sam = mlp->w[l];
    g.sak = 1;
    g.B = mlp->wb + mlp->woff[l];
    g.sbk = mlp->w[l + 1];
    g.sbn = 1;
    g.M = B;
    g.N = mlp->w[l + 1];
    g.K = mlp->w[l];
    g.bias = mlp->p + mlp->boff[l];
    if (l + 1 < mlp->L) {
      g.Z = mlp->Z[l];
      g.H = mlp->H[l];
      if ((rc = gemm<EPI_GELU>(mlp, s, g))) return rc;
      h = mlp->H[l];
    } else {
      g.F = y;
      if ((rc = gemm<EPI_LINEAR>(mlp, s, g))) return rc;
    }
  }
  mlp->x = (const __bf16*)x;
  mlp->last_batch = B;
  return DISORT_OK;
}

// Gradients of the parameters for output gradient dy [B x w_L] (BF16 bits,
// device) of the last forward, into the packed FP32 gradient vector.
int disort_mlp_backward(disort_ctx* ctx, disort_mlp* mlp, const uint16_t* dy) {
  if (!ctx || !mlp || !dy || !mlp->x) return mfail(mlp, DISORT_E_INVALID, "backward needs a preceding forward");
  hipStream_t s = stream_of(ctx);
  const int B = mlp->last_batch;
  const __bf16* dz = (const __bf16*)dy;
  int rc;
  for (int l = mlp->L - 1; l >= 0; --l) {
    const __bf16* hin = (l == 0) ? mlp->x : mlp->H[l - 1];
    GemmArgs gw{};  // dW_l = h_l^T dz
    gw.A = hin;
    gw.sam = 1;
    gw.sak = mlp->w[l];
    gw.B = dz;
    gw.sbk = mlp->w[l + 1];
    gw.sbn = 1;
    gw.M = mlp->w[l];
    gw.N = mlp->w[l + 1];
    gw.K = B;
    gw.F = mlp->grad + mlp->woff[l];
    if ((rc = gemm<EPI_WGRAD>(mlp, s, gw))) return rc;
    hipLaunchKernelGGL(k_colsum, dim3(blocks(mlp->w[l + 1])), dim3(256), 0, s, dz, B, mlp->w[l + 1],
                       mlp->grad + mlp->boff[l]);
    MCHK(mlp, hipGetLastError());
    if (l > 0) {  // dz_{l-1} = (dz W_l^T) * GELU'(z_{l-1})
      __bf16* next = (dz == mlp->dz0) ? mlp->dz1 : mlp->dz0;
      GemmArgs gd{};
      gd.A = dz;
      gd.sam = mlp->w[l + 1];
      gd.sak = 1;
      gd.B = mlp->wb + mlp->woff[l];
      gd.sbk = 1;
      gd.sbn = mlp->w[l + 1];
      gd.M = B;
      gd.N = mlp->w[l];
      gd.K = mlp->w[l + 1];
      gd.H = next;
      gd.Zin = mlp->Z[l - 1];
      if ((rc = gemm<EPI_DGELU>(mlp, s, gd))) return rc;
      dz = next;
    }
  }
  return DISORT_OK;
}

// One Adam step (t >= 1) on the current gradients.
int disort_mlp_adam(disort_ctx* ctx, disort_mlp* mlp, float lr, float beta1, float beta2, float eps, int t) {
  if (!ctx || !mlp || t < 1) return DISORT_E_INVALID;
  hipStream_t s = stream_of(ctx);
  const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3(blocks(mlp->nparam)), dim3(256), 0, s, mlp->p, mlp->grad, mlp->m, mlp->v, mlp->wb,
                     mlp->nparam, lr, beta1, beta2, eps, c1, c2);
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
  hipLaunchKernelGGL(k_mse, dim3(blocks((int64_t)B * N)), dim3(256), 0, s, mlp->y, tgt, mask_dev, B, N, scale, mlp->dy,
                     loss_dev);
  MCHK(mlp, hipGetLastError());
  if ((rc = disort_mlp_backward(ctx, mlp, (const uint16_t*)mlp->dy))) return rc;
  return disort_mlp_adam(ctx, mlp, lr, beta1, beta2, eps, t);
}

}  // extern "C"
