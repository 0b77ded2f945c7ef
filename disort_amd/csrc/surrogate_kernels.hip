// Column-level kernels of the flux surrogate (no counterpart in the
// reference): input features and training targets straight from the
// device-resident flux batch, decoding of the network outputs into fluxes
// with the ensemble spread used by the gate, and the gather / scatter that
// routes gated-out columns through the plain solver.
//
// Columns are uniform: n columns of nlyr layers, moment stride pstride,
// arrays laid out as disort_flux_batch with lay_off[i] = i*nlyr.
// Features of column i (4*nlyr + 2 of them, then zero padding to w0):
//   [l]          log10(delta_tau_l)
//   [nlyr + l]   single_scat_albedo_l
//   [2nlyr + l]  g_l = pmom_l[1]
//   [3nlyr + l]  layer-mean temperature (T_l + T_{l+1}) / 2
//   [4nlyr]      mu0,  [4nlyr + 1] surface albedo
// Targets (3*nlev, then zero padding to wL), in units of the incident beam
// mu0*F0 at the nlev = nlyr+1 levels:
//   [v]          (flux_down - flux_up) / (mu0 F0)   diffuse net flux
//   [nlev + v]   (flux_down + flux_up) / (mu0 F0)
//   [2nlev + v]  flux_direct_beam / (mu0 F0)
// (the network's output space: the loss then weighs the net flux directly).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/disort_amd.h"

namespace {

unsigned blocks(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

__global__ void k_features(int64_t n, int64_t n_pad, int L, int P, const double* dtau, const double* ssa,
                           const double* pmom, const double* temper, const double* scal, const float* shift,
                           const float* scale, const float* lo, const float* hi, int w0, __bf16* x, int32_t* ood) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pad) return;
  __bf16* xr = x + i * w0;
  if (i >= n) {
    for (int k = 0; k < w0; ++k) xr[k] = (__bf16)0.0f;
    return;
  }
  const int nf = 4 * L + 2;
  int bad = 0;
  for (int k = 0; k < w0; ++k) {
    double f = 0.0;
    if (k < L) f = log10(dtau[i * L + k]);
    else if (k < 2 * L) f = ssa[i * L + (k - L)];
    else if (k < 3 * L) f = pmom[(i * L + (k - 2 * L)) * P + 1];
    else if (k < 4 * L) f = 0.5 * (temper[i * (L + 1) + (k - 3 * L)] + temper[i * (L + 1) + (k - 3 * L) + 1]);
    else if (k == 4 * L) f = scal[i * DISORT_NSCAL + DISORT_S_BEAM_MU];
    else if (k == 4 * L + 1) f = scal[i * DISORT_NSCAL + DISORT_S_ALBEDO];
    if (k < nf) {
      const float ff = (float)f;
      bad |= !(ff >= lo[k] && ff <= hi[k]);
      xr[k] = (__bf16)((ff - shift[k]) * scale[k]);
    } else {
      xr[k] = (__bf16)0.0f;
    }
  }
  if (ood) ood[i] = bad;
}

__device__ __forceinline__ double beam_unit(const double* scal, int64_t i) {
  const double u = scal[i * DISORT_NSCAL + DISORT_S_BEAM_MU] * scal[i * DISORT_NSCAL + DISORT_S_BEAM_FLUX];
  return u > 1e-300 ? u : 1.0;
}

__global__ void k_targets(int64_t n, int64_t n_pad, int nlev, const double* out, int64_t tv, const double* scal,
                          const float* shift, const float* scale, int wL, float* t) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_pad * wL) return;
  const int64_t i = e / wL;
  const int k = (int)(e % wL);
  float v = 0.f;
  if (i < n && k < 3 * nlev) {
    const int q = k / nlev, lv = k % nlev;
    const double up = out[DISORT_O_FLUX_UP * tv + i * nlev + lv];
    const double dn = out[DISORT_O_FLUX_DOWN * tv + i * nlev + lv];
    const double dir = out[DISORT_O_FLUX_DIRECT_BEAM * tv + i * nlev + lv];
    const double u = beam_unit(scal, i);
    const double x = (q == 0) ? (dn - up) / u : (q == 1 ? (dn + up) / u : dir / u);
    v = ((float)x - shift[k]) * scale[k];
  }
  t[e] = v;
}

// Network outputs (standardised) -> fluxes [n][3][nlev] (up, down, direct),
// averaged over the ensemble members; the direct beam is replaced by its
// closed form mu0 F0 exp(-tau/mu0) on the unscaled optical depth, which is
// what the plain solver returns for it (src/FluxSolver.cpp:1187-1190).
// spread = max over levels of |Fnet_1 - Fnet_2| / max(c_local |Fnet|, c_floor mu0 F0)
// (c_local = 1, c_floor = 1e-3: relative to the local net flux with a floor;
// c_local = 0, c_floor = 1: relative to the incident beam).
__global__ void k_decode(int64_t n, int L, int wL, const float* y1, const float* y2, const float* shift,
                         const float* scale, const double* dtau, const double* scal, double c_local, double c_floor,
                         double* flux, float* spread) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int nlev = L + 1;
  const double mu0 = scal[i * DISORT_NSCAL + DISORT_S_BEAM_MU], F0 = scal[i * DISORT_NSCAL + DISORT_S_BEAM_FLUX];
  const double u = beam_unit(scal, i);
  const double floor = c_floor * u;
  double tau = 0.0, worst = 0.0;
  for (int lv = 0; lv < nlev; ++lv) {
    if (lv > 0) tau += dtau[i * L + lv - 1];
    const double dir = (F0 > 0.0) ? mu0 * F0 * exp(-tau / mu0) : 0.0;
    const int ka = lv, ks = nlev + lv;
    const double a1 = ((double)y1[i * wL + ka] / scale[ka] + shift[ka]) * u;  // diffuse net
    const double s1 = ((double)y1[i * wL + ks] / scale[ks] + shift[ks]) * u;  // diffuse sum
    double a = a1, sm = s1;
    if (y2) {
      const double a2 = ((double)y2[i * wL + ka] / scale[ka] + shift[ka]) * u;
      const double s2 = ((double)y2[i * wL + ks] / scale[ks] + shift[ks]) * u;
      a = 0.5 * (a1 + a2);
      sm = 0.5 * (s1 + s2);
      const double net = c_local * fabs(dir + a);
      const double den = net > floor ? net : floor;
      const double sp = fabs(a1 - a2) / den;
      worst = sp > worst ? sp : worst;
    }
    double* fr = flux + i * 3 * nlev;
    fr[lv] = 0.5 * (sm - a);
    fr[nlev + lv] = 0.5 * (sm + a);
    fr[2 * nlev + lv] = dir;
  }
  spread[i] = (float)worst;
}

// dst column k <- src column idx[k] (uniform layout).
__global__ void k_gather(int64_t m, const int32_t* idx, int L, int P, const double* dtau, const double* ssa,
                         const double* pmom, const double* temper, const double* scal, double* odtau, double* ossa,
                         double* opmom, double* otemper, double* oscal) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int64_t i = idx[k];
  for (int l = 0; l < L; ++l) {
    odtau[k * L + l] = dtau[i * L + l];
    ossa[k * L + l] = ssa[i * L + l];
  }
  for (int q = 0; q < L * P; ++q) opmom[k * L * P + q] = pmom[i * L * P + q];
  if (temper)
    for (int l = 0; l <= L; ++l) otemper[k * (L + 1) + l] = temper[i * (L + 1) + l];
  for (int q = 0; q < DISORT_NSCAL; ++q) oscal[k * DISORT_NSCAL + q] = scal[i * DISORT_NSCAL + q];
}

// flux rows idx[k] <- plain-solver outputs of sub-batch column k.
__global__ void k_scatter(int64_t m, const int32_t* idx, int nlev, const double* out, int64_t tv, double* flux) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m * nlev) return;
  const int64_t k = e / nlev;
  const int lv = (int)(e % nlev);
  double* fr = flux + (int64_t)idx[k] * 3 * nlev;
  fr[lv] = out[DISORT_O_FLUX_UP * tv + k * nlev + lv];
  fr[nlev + lv] = out[DISORT_O_FLUX_DOWN * tv + k * nlev + lv];
  fr[2 * nlev + lv] = out[DISORT_O_FLUX_DIRECT_BEAM * tv + k * nlev + lv];
}

// Rows of a row-major matrix (elem_bytes 2 or 4): dst[k] <- src[idx[k]].
__global__ void k_rows(int64_t m, const int32_t* idx, int64_t row_elems, int elem_bytes, const void* src, void* dst) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m * row_elems) return;
  const int64_t k = e / row_elems, c = e % row_elems;
  const int64_t s = (int64_t)idx[k] * row_elems + c;
  if (elem_bytes == 2) ((uint16_t*)dst)[e] = ((const uint16_t*)src)[s];
  else ((uint32_t*)dst)[e] = ((const uint32_t*)src)[s];
}

}  // namespace

extern "C" {

#define SURR_LAUNCH_CHECK()                                          \
  do {                                                               \
    if (hipGetLastError() != hipSuccess) return DISORT_E_HIP;        \
  } while (0)

int disort_surrogate_features(disort_ctx* ctx, int64_t n, int64_t n_pad, int nlyr, int pstride,
                              const disort_flux_batch* b, const float* shift, const float* scale, const float* lo,
                              const float* hi, int w0, uint16_t* x, int32_t* ood) {
  if (!ctx || !b || n < 0 || n_pad < n || nlyr <= 0 || pstride < 2 || w0 < 4 * nlyr + 2) return DISORT_E_INVALID;
  if (!b->temper) return DISORT_E_INVALID;
  hipLaunchKernelGGL(k_features, dim3(blocks(n_pad)), dim3(256), 0, (hipStream_t)disort_ctx_stream(ctx), n, n_pad, nlyr,
                     pstride, b->dtau, b->ssa, b->pmom, b->temper, b->scal, shift, scale, lo, hi, w0, (__bf16*)x, ood);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_surrogate_targets(disort_ctx* ctx, int64_t n, int64_t n_pad, int nlev, const disort_flux_batch* b,
                             const float* shift, const float* scale, int wL, float* t) {
  if (!ctx || !b || !b->out || n < 0 || n_pad < n || wL < 3 * nlev) return DISORT_E_INVALID;
  hipLaunchKernelGGL(k_targets, dim3(blocks(n_pad * wL)), dim3(256), 0, (hipStream_t)disort_ctx_stream(ctx), n, n_pad,
                     nlev, b->out, b->total_levels, b->scal, shift, scale, wL, t);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_surrogate_decode(disort_ctx* ctx, int64_t n, int nlyr, int wL, const float* y1, const float* y2,
                            const float* shift, const float* scale, const disort_flux_batch* b, double c_local,
                            double c_floor, double* flux, float* spread) {
  if (!ctx || !b || !y1 || !flux || !spread || wL < 3 * (nlyr + 1) || !(c_floor > 0.0)) return DISORT_E_INVALID;
  hipLaunchKernelGGL(k_decode, dim3(blocks(n)), dim3(256), 0, (hipStream_t)disort_ctx_stream(ctx), n, nlyr, wL, y1, y2,
                     shift, scale, b->dtau, b->scal, c_local, c_floor, flux, spread);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_surrogate_gather(disort_ctx* ctx, int64_t m, const int32_t* idx, int nlyr, int pstride,
                            const disort_flux_batch* src, disort_flux_batch* dst) {
  if (!ctx || !idx || !src || !dst || m < 0) return DISORT_E_INVALID;
  if (m == 0) return DISORT_OK;
  hipLaunchKernelGGL(k_gather, dim3(blocks(m, 64)), dim3(64), 0, (hipStream_t)disort_ctx_stream(ctx), m, idx, nlyr,
                     pstride, src->dtau, src->ssa, src->pmom, src->temper, src->scal, (double*)dst->dtau,
                     (double*)dst->ssa, (double*)dst->pmom, (double*)dst->temper, (double*)dst->scal);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_surrogate_scatter(disort_ctx* ctx, int64_t m, const int32_t* idx, int nlev, const double* out,
                             int64_t total_levels, double* flux) {
  if (!ctx || !idx || !out || !flux || m < 0) return DISORT_E_INVALID;
  if (m == 0) return DISORT_OK;
  hipLaunchKernelGGL(k_scatter, dim3(blocks(m * nlev)), dim3(256), 0, (hipStream_t)disort_ctx_stream(ctx), m, idx, nlev,
                     out, total_levels, flux);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_rows_gather(disort_ctx* ctx, int64_t m, const int32_t* idx, int64_t row_elems, int elem_bytes,
                       const void* src, void* dst) {
  if (!ctx || !idx || !src || !dst || m < 0 || (elem_bytes != 2 && elem_bytes != 4)) return DISORT_E_INVALID;
  if (m == 0) return DISORT_OK;
  hipLaunchKernelGGL(k_rows, dim3(blocks(m * row_elems)), dim3(256), 0, (hipStream_t)disort_ctx_stream(ctx), m, idx,
                     row_elems, elem_bytes, src, dst);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

}  // extern "C"
