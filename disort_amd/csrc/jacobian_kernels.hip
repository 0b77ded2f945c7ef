// Device side of the device-resident temperature Jacobians
// (disort_jac_plan in capi.cpp; the compute_temperature_jacobian option of
// DisortFluxSolver<N>::solve, src/FluxSolver.cpp:1294-1497).
//
// With the optical properties fixed, every flux output is affine in the
// level Planck values B_k and in B(T_bottom).  A chunk of uniform columns
// is expanded into one solver instance per (column, degree of freedom):
// the column's medium with a unit Planck value at that degree of freedom,
// zero elsewhere, and no beam, isotropic or top emission.  The expanded
// solves' outputs times dB/dT at that degree of freedom are the Jacobian
// columns, written in the layout of disort_flux_jacobian:
// jac[i][field][level lu][dof k], fields flux_up, flux_down,
// mean_intensity, mean_intensity_down, mean_intensity_up,
// flux_tau_divergence; k = 0..L the level temperatures, L+1 the surface.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/disort_amd.h"
#include "flux_layout.hpp"
#include "planck.hpp"

namespace {

// One workgroup per expanded instance e of the chunk starting at column c0:
// column min(c0 + e / (L+2), n-1) (a last partial chunk repeats column
// n-1), degree of freedom e % (L+2).  Writes the expanded batch arrays, the
// Planck override (level values [e*(L+1) + v], then top and bottom pairs
// after the TV = ne*(L+1) level entries) and dB/dT of the degree of freedom.
__global__ __launch_bounds__(256) void k_jac_expand(int64_t n, int64_t c0, int64_t ne, int L, int P,
                                                    const double* __restrict__ dtau, const double* __restrict__ ssa,
                                                    const double* __restrict__ pmom,
                                                    const double* __restrict__ temper,
                                                    const double* __restrict__ scal, double* xdtau, double* xssa,
                                                    double* xpmom, double* xtemper, double* xscal, double* ovr,
                                                    double* dbdt) {
  const int64_t e = blockIdx.x;
  if (e >= ne) return;
  const int D = L + 2, t = threadIdx.x;
  int64_t i = c0 + e / D;
  if (i > n - 1) i = n - 1;
  const int k = (int)(e % D);
  for (int j = t; j < L; j += blockDim.x) {
    xdtau[e * L + j] = dtau[i * L + j];
    xssa[e * L + j] = ssa[i * L + j];
  }
  const int64_t mp = (int64_t)L * P;
  for (int64_t j = t; j < mp; j += blockDim.x) xpmom[e * mp + j] = pmom[i * mp + j];
  for (int j = t; j <= L; j += blockDim.x) {
    xtemper[e * (L + 1) + j] = temper[i * (L + 1) + j];
    ovr[e * (L + 1) + j] = (j == k) ? 1.0 : 0.0;
  }
  const double* sc = scal + i * DISORT_NSCAL;
  if (t < DISORT_NSCAL) {
    const bool src = t == DISORT_S_BEAM_FLUX || t == DISORT_S_ISO_TOP || t == DISORT_S_ISO_BOT;
    xscal[e * DISORT_NSCAL + t] = src ? 0.0 : sc[t];
  }
  if (t == 0) {
    const int64_t TV = ne * (L + 1);
    ovr[TV + 2 * e] = 0.0;
    ovr[TV + 2 * e + 1] = (k == L + 1) ? 1.0 : 0.0;
    const double T = (k <= L) ? temper[i * (L + 1) + k] : sc[DISORT_S_TEMP_BOT];
    dbdt[e] = dsa::planck_band_deriv(sc[DISORT_S_WN_LO], sc[DISORT_S_WN_HI], T);
  }
}

// jac rows of columns c0 .. c0+cols-1: one thread per Jacobian entry,
// entries of one column contiguous (coalesced writes).
__global__ void k_jac_scale(int64_t cols, int64_t c0, int L, const double* __restrict__ out, int64_t TV,
                            const double* __restrict__ dbdt, double* jac) {
  const int D = L + 2;
  const int64_t per = (int64_t)(L + 1) * D;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= cols * 6 * per) return;
  const int64_t ci = q / (6 * per);
  const int64_t r = q - ci * 6 * per;
  const int f = (int)(r / per);
  const int64_t r2 = r - f * per;
  const int lu = (int)(r2 / D), k = (int)(r2 % D);
  const int field = (f == 0)   ? DISORT_O_FLUX_UP
                    : (f == 1) ? DISORT_O_FLUX_DOWN
                    : (f == 2) ? DISORT_O_MEAN_INTENSITY
                    : (f == 3) ? DISORT_O_MEAN_INTENSITY_DOWN
                    : (f == 4) ? DISORT_O_MEAN_INTENSITY_UP
                               : DISORT_O_FLUX_TAU_DIVERGENCE;
  const int64_t e = ci * D + k;
  jac[(c0 + ci) * 6 * per + r] = dbdt[e] * out[(int64_t)field * TV + e * (L + 1) + lu];
}

// Column status of the chunk: the first nonzero status of its expanded
// instances (one wavefront per column, lanes over degrees of freedom).
__global__ void k_jac_status(int64_t cols, int64_t c0, int L, const int32_t* __restrict__ inst_i, int32_t* status) {
  const int64_t ci = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (ci >= cols) return;
  const int D = L + 2;
  int32_t s = 0;
  for (int k = lane; k < D; k += 64) {
    const int32_t v = inst_i[(ci * D + k) * dsa::kInstI + dsa::II_STATUS];
    if (s == 0) s = v;
  }
  for (int off = 32; off >= 1; off >>= 1) {
    const int32_t o = __shfl_xor(s, off);
    s = s != 0 ? s : o;
  }
  if (lane == 0) status[c0 + ci] = s;
}

unsigned blocks(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

namespace dsa {

bool launch_jac_expand(int64_t n, int64_t c0, int64_t ne, int L, int P, const disort_flux_batch* b, double* xdtau,
                       double* xssa, double* xpmom, double* xtemper, double* xscal, double* ovr, double* dbdt,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_jac_expand, dim3((unsigned)ne), dim3(256), 0, st, n, c0, ne, L, P, b->dtau, b->ssa, b->pmom,
                     b->temper, b->scal, xdtau, xssa, xpmom, xtemper, xscal, ovr, dbdt);
  return hipGetLastError() == hipSuccess;
}

bool launch_jac_scale(int64_t cols, int64_t c0, int L, const double* out, int64_t TV, const double* dbdt,
                      double* jac, const int32_t* inst_i, int32_t* status, hipStream_t st) {
  const int64_t per = (int64_t)(L + 1) * (L + 2);
  hipLaunchKernelGGL(k_jac_scale, dim3(blocks(cols * 6 * per)), dim3(256), 0, st, cols, c0, L, out, TV, dbdt, jac);
  hipLaunchKernelGGL(k_jac_status, dim3(blocks(cols, 4)), dim3(256), 0, st, cols, c0, L, inst_i, status);
  return hipGetLastError() == hipSuccess;
}

}  // namespace dsa
