// Column-level kernels of the flux surrogate (no counterpart in the
// reference): input features and training targets straight from the
// device-resident flux batch, decoding of the network outputs into fluxes
// with the ensemble spread used by the gate, and the gather / scatter that
// routes gated-out columns through the plain solver.
//
// Columns are uniform: n columns of nlyr layers, moment stride pstride,
// arrays laid out as disort_flux_batch with lay_off[i] = i*nlyr.
// Features of column i (4*nlyr + 2 of them, with w0 >= 7*nlyr + 5 another
// 3*(nlyr+1) level features, then zero padding to w0):
//   [l]          log10(delta_tau_l)
//   [nlyr + l]   single_scat_albedo_l
//   [2nlyr + l]  g_l = pmom_l[1]
//   [3nlyr + l]  layer-mean temperature (T_l + T_{l+1}) / 2
//   [4nlyr]      mu0,  [4nlyr + 1] surface albedo
//   [4nlyr + 2 + v]              log10(tau_v + 1e-3), tau_v the optical depth of level v
//   [4nlyr + 2 + (nlyr+1) + v]   exp(-tau_v / mu0), the direct-beam transmission
//   [4nlyr + 2 + 2(nlyr+1) + v]  log10(sum_{l<v} (1 - omega_l) dtau_l + 1e-3)
// and, with w0 >= 8*nlyr + 7, the thermal source in units of the beam:
//   [4nlyr + 2 + 3(nlyr+1) + v]  B(T_v) / (mu0 F0), B the band Planck function
//   [4nlyr + 2 + 4(nlyr+1)]      (1 - albedo) B(T_bottom) / (mu0 F0)
// Targets (3*nlev, then zero padding to wL), in units of the incident beam
// mu0*F0 at the nlev = nlyr+1 levels:
//   [v]          (flux_down - flux_up) / (mu0 F0)   diffuse net flux
//   [nlev + v]   (flux_down + flux_up) / (mu0 F0)
//   [2nlev + v]  flux_direct_beam / (mu0 F0)
// (the network's output space: the loss then weighs the net flux directly).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "../../include/disort_amd.h"
#include "planck.hpp"

namespace {

unsigned blocks(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// Row length of the two-stream baseline array ts: U and D at the nlyr+1
// levels, the level and surface Planck values, the layer asymmetry factors.
__host__ __device__ constexpr int ts_row(int L) { return 3 * (L + 1) + 1 + L; }

// Features end up in BF16 (8 significant bits), so their logarithms and
// exponentials are taken in single precision (about 1e-7 relative).
__device__ __forceinline__ double flog10(double x) { return (double)log10f((float)x); }
__device__ __forceinline__ double fexp(double x) { return (double)expf((float)x); }

__device__ __forceinline__ double beam_unit(const double* scal, int64_t i) {
  const double u = scal[i * DISORT_NSCAL + DISORT_S_BEAM_MU] * scal[i * DISORT_NSCAL + DISORT_S_BEAM_FLUX];
  return u > 1e-300 ? u : 1.0;
}

// One wavefront per column (four per workgroup): lane v first forms the
// values of level v -- optical depth and absorption optical depth summed
// from the top in layer order, beam transmission, band Planck value -- into
// LDS, then the lanes write the column's features, 64 at a time.
__global__ __launch_bounds__(256) void k_features(int64_t n, int64_t n_pad, int L, int P, const double* dtau,
                                                  const double* ssa, const double* pmom, const double* temper,
                                                  const double* scal, const float* shift, const float* scale,
                                                  const float* lo, const float* hi, int w0, __bf16* x, int32_t* ood,
                                                  const double* ts) {
  __shared__ double lev[4][4][64];  // [column][tau, absorption tau, transmission, B/(mu0 F0)][level]
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + wv;
  if (i >= n_pad) return;
  __bf16* xr = x + i * w0;
  if (i >= n) {
    for (int k = lane; k < w0; k += 64) xr[k] = (__bf16)0.0f;
    return;
  }
  const int base = 4 * L + 2, nlev = L + 1;
  const int ext = (w0 >= base + 4 * nlev + 1) ? 4 * nlev + 1 : (w0 >= base + 3 * nlev) ? 3 * nlev : 0;
  // Two-stream baseline block (net and sum of its diffuse fluxes in beam units).
  const int tsb = base + 4 * nlev + 1;
  const bool with_ts = ts && ext == 4 * nlev + 1 && w0 >= tsb + 2 * nlev;
  const int nf = base + ext + (with_ts ? 2 * nlev : 0);
  const double* tsr = with_ts ? ts + i * ts_row(L) : nullptr;
  const double* sc = scal + i * DISORT_NSCAL;
  const double mu0 = sc[DISORT_S_BEAM_MU];
  const double u = beam_unit(scal, i);
  double (*lv)[64] = lev[wv];
  // Optical and absorption depth of level v = lane: exclusive prefix sums
  // over the layers by a wavefront scan (L + 1 < 64).
  double dsum = 0.0, asum = 0.0;
  if (lane < L) {
    dsum = dtau[i * L + lane];
    asum = (1.0 - ssa[i * L + lane]) * dsum;
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double td = __shfl_up(dsum, off), ta = __shfl_up(asum, off);
    if (lane >= off) {
      dsum += td;
      asum += ta;
    }
  }
  const double tau_ex = __shfl_up(dsum, 1), tab_ex = __shfl_up(asum, 1);
  {
    const int v = lane;  // v == nlev: the surface (Planck term only)
    const double tau = (lane > 0) ? tau_ex : 0.0, tab = (lane > 0) ? tab_ex : 0.0;
    if (ext > 0 && v < nlev) {
      lv[0][v] = flog10(tau + 1e-3);
      lv[1][v] = flog10(tab + 1e-3);
      lv[2][v] = fexp(-tau / mu0);
    }
    if (ext > 3 * nlev && v <= nlev) {
      double bb;
      if (with_ts) {
        bb = tsr[2 * nlev + v];  // the baseline's level Planck values
      } else {
        const double t = (v < nlev) ? temper[i * nlev + v] : sc[DISORT_S_TEMP_BOT];
        bb = dsa::planck_band(sc[DISORT_S_WN_LO], sc[DISORT_S_WN_HI], t);
      }
      lv[3][v] = (v < nlev ? bb : (1.0 - sc[DISORT_S_ALBEDO]) * bb) / u;
    }
  }
  // Only this wavefront reads its LDS rows: order the LDS writes before the
  // reads within the wavefront.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int bad = 0;
  for (int k = lane; k < w0; k += 64) {
    double f = 0.0;
    if (k < L) f = flog10(dtau[i * L + k]);
    else if (k < 2 * L) f = ssa[i * L + (k - L)];
    else if (k < 3 * L) f = with_ts ? tsr[3 * nlev + 1 + (k - 2 * L)] : pmom[(i * L + (k - 2 * L)) * P + 1];
    else if (k < 4 * L) f = 0.5 * (temper[i * nlev + (k - 3 * L)] + temper[i * nlev + (k - 3 * L) + 1]);
    else if (k == 4 * L) f = mu0;
    else if (k == 4 * L + 1) f = sc[DISORT_S_ALBEDO];
    else if (k >= tsb && k < nf) {
      const int v = (k - tsb) % nlev;
      const double up = tsr[v], dn = tsr[nlev + v];
      f = ((k - tsb) < nlev ? dn - up : dn + up) / u;
    } else if (k < nf) {
      const int blk = (k - base) / nlev, v = (k - base) % nlev;
      // blocks: optical depth, transmission, absorption depth, Planck (the last has nlev + 1 entries)
      f = (blk == 0) ? lv[0][v] : (blk == 1) ? lv[2][v] : (blk == 2) ? lv[1][v] : lv[3][k - base - 3 * nlev];
    }
    if (k < nf) {
      const float ff = (float)f;
      bad |= !(ff >= lo[k] && ff <= hi[k]);
      xr[k] = (__bf16)((ff - shift[k]) * scale[k]);
    } else {
      xr[k] = (__bf16)0.0f;
    }
  }
  const unsigned long long any_bad = __ballot(bad);
  if (ood && lane == 0) ood[i] = any_bad ? 1 : 0;
}

__global__ void k_targets(int64_t n, int64_t n_pad, int nlev, const double* out, int64_t tv, const double* scal,
                          const float* shift, const float* scale, int wL, float* t, const double* ts) {
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
    double x = (q == 0) ? (dn - up) / u : (q == 1 ? (dn + up) / u : dir / u);
    if (ts && q < 2) {
      const double bu = ts[i * ts_row(nlev - 1) + lv], bd = ts[i * ts_row(nlev - 1) + nlev + lv];
      x -= (q == 0 ? bd - bu : bd + bu) / u;  // residual over the two-stream baseline
    }
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
constexpr int kMaxMembers = 8;
struct Members {
  const float* y[kMaxMembers];
  int k;
};

__global__ __launch_bounds__(256) void k_decode(int64_t n, int L, int wL, Members ms, const float* shift,
                                                const float* scale, const double* dtau, const double* scal,
                                                double c_local, double c_floor, double* flux, float* spread,
                                                const double* ts) {
  // One wavefront per column, lane = level (nlev <= 64).
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const int nlev = L + 1;
  const double mu0 = scal[i * DISORT_NSCAL + DISORT_S_BEAM_MU], F0 = scal[i * DISORT_NSCAL + DISORT_S_BEAM_FLUX];
  const double u = beam_unit(scal, i);
  const double floor = c_floor * u;
  double worst = 0.0;
  __shared__ double lay[4][64];
  double* ly = lay[threadIdx.x >> 6];
  if (lane < L) ly[lane] = dtau[i * L + lane];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int lv = lane;
  if (lv < nlev) {
    double tau = 0.0;
    for (int l = 0; l < lv; ++l) tau += ly[l];  // the solver's summation order
    const double dir = (F0 > 0.0) ? mu0 * F0 * exp(-tau / mu0) : 0.0;
    const int ka = lv, ks = nlev + lv;
    // Two-stream baseline (zero without ts): the network predicts the residual.
    const double* tr = ts ? ts + i * ts_row(nlev - 1) : nullptr;
    const double bn = ts ? tr[nlev + lv] - tr[lv] : 0.0;
    const double bs = ts ? tr[nlev + lv] + tr[lv] : 0.0;
    // Members' diffuse net and sum; their mean is served, and twice the
    // standard deviation of the net flux (|a_1 - a_2| for two members) is
    // the spread.
    double am[kMaxMembers];
    double asum = 0.0, ssum = 0.0;
#pragma unroll
    for (int m = 0; m < kMaxMembers; ++m) {
      if (m >= ms.k) break;
      const float* y = ms.y[m];
      am[m] = ((double)y[i * wL + ka] / scale[ka] + shift[ka]) * u + bn;
      asum += am[m];
      ssum += ((double)y[i * wL + ks] / scale[ks] + shift[ks]) * u + bs;
    }
    const double a = asum / ms.k, sm = ssum / ms.k;
    if (ms.k > 1) {
      double var = 0.0;
#pragma unroll
      for (int m = 0; m < kMaxMembers; ++m) {
        if (m >= ms.k) break;
        var += (am[m] - a) * (am[m] - a);
      }
      const double net = c_local * fabs(dir + a);
      const double den = net > floor ? net : floor;
      worst = 2.0 * sqrt(var / ms.k) / den;
    }
    // A non-finite output fails the gate (one member included).
    if (!(worst == worst) || !isfinite(a) || !isfinite(sm)) worst = __builtin_inf();
    double* fr = flux + i * 3 * nlev;
    fr[lv] = 0.5 * (sm - a);
    fr[nlev + lv] = 0.5 * (sm + a);
    fr[2 * nlev + lv] = dir;
  }
  for (int off = 32; off >= 1; off >>= 1) worst = fmax(worst, __shfl_xor(worst, off));
  if (lane == 0) spread[i] = (float)worst;
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

// Rows of a row-major matrix (elem_bytes 2, 4 or 8): dst[k] <- src[idx[k]];
// consecutive threads copy consecutive elements of a row (coalesced).
__global__ void k_rows(int64_t m, const int32_t* idx, int64_t row_elems, int elem_bytes, const void* src, void* dst) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m * row_elems) return;
  const int64_t k = e / row_elems, c = e % row_elems;
  const int64_t s = (int64_t)idx[k] * row_elems + c;
  if (elem_bytes == 2) ((uint16_t*)dst)[e] = ((const uint16_t*)src)[s];
  else if (elem_bytes == 4) ((uint32_t*)dst)[e] = ((const uint32_t*)src)[s];
  else ((uint64_t*)dst)[e] = ((const uint64_t*)src)[s];
}

// Counter-based uniform draw in [0, 1) for column i of a serve call
// (SplitMix64 finaliser of seed + i): the sampled re-solve needs no host RNG.
__device__ __forceinline__ double unit_draw(uint64_t seed, int64_t i) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void k_zero_i32(int32_t* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// Gate decision per column: flags bit 0 routed to plain DISORT (out of
// distribution, spread above tau or not a number, sampled, or route_all), bit 1
// sampled, bit 2 passes the gate.  Routed indices are appended to idx (one
// atomic per wavefront).
__global__ void k_route(int64_t n, const float* spread, const int32_t* ood, float tau, double frac, uint64_t seed,
                        int route_all, int32_t* flags, int32_t* idx, int32_t* count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = i < n;
  bool routed = false;
  if (in) {
    const bool sampled = unit_draw(seed, i) < frac;
    const bool pass = !(ood && ood[i] != 0) && spread[i] <= tau;
    routed = route_all || sampled || !pass;
    flags[i] = (routed ? 1 : 0) | (sampled ? 2 : 0) | (pass ? 4 : 0);
  }
  const unsigned long long bal = __ballot(routed);
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (lane == 0 && bal) base = atomicAdd(count, __popcll(bal));
  base = __shfl(base, 0);
  if (routed) idx[base + __popcll(bal & ((1ull << lane) - 1ull))] = (int32_t)i;
}

// idx[m .. mpad) = idx[0]: padding of a routed list to a cached sub-batch
// size (the copies solve the same column again and scatter the same values).
__global__ void k_pad_idx(int64_t m, int64_t mpad, int32_t* idx) {
  const int64_t k = m + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < mpad) idx[k] = idx[0];
}

// Online error monitor: for every routed column that was sampled although
// it passes the gate, the surrogate's net-flux error against the plain solve (max over levels of
// |dF_net| / max(c_local |F_net|, c_floor mu0 F0)); counters[0] += sampled
// columns, counters[1] += those above tol.
__global__ void k_monitor(int64_t m, const int32_t* idx, const int32_t* flags, int nlev, const double* out, int64_t tv,
                          const double* pred, const double* scal, double c_local, double c_floor, double tol,
                          int32_t* counters) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool smp = false, bad = false;
  if (k < m) {
    const int64_t i = idx[k];
    smp = (flags[i] & 6) == 6;
    if (smp) {
      const double u = beam_unit(scal, i), floor = c_floor * u;
      const double* fp = pred + i * 3 * nlev;
      double worst = 0.0;
      for (int lv = 0; lv < nlev; ++lv) {
        const double rnet = out[DISORT_O_FLUX_DIRECT_BEAM * tv + k * nlev + lv] +
                            out[DISORT_O_FLUX_DOWN * tv + k * nlev + lv] - out[DISORT_O_FLUX_UP * tv + k * nlev + lv];
        const double pnet = fp[2 * nlev + lv] + fp[nlev + lv] - fp[lv];
        const double den = fmax(c_local * fabs(rnet), floor);
        worst = fmax(worst, fabs(pnet - rnet) / den);
      }
      bad = !(worst <= tol);
    }
  }
  const unsigned long long bs = __ballot(smp), bb = __ballot(bad);
  if ((threadIdx.x & 63) == 0) {
    if (bs) atomicAdd(&counters[0], __popcll(bs));
    if (bb) atomicAdd(&counters[1], __popcll(bb));
  }
}

// ---------------------------------------------------------------------------
// Two-stream baseline of the flux surrogate: per column (one thread each),
// delta-Eddington scaling (f = g^2), quadrature coefficients
//   gamma1 = sqrt3 (2 - w(1+g))/2, gamma2 = sqrt3 w (1-g)/2,
//   gamma3 = (1 - sqrt3 g mu0)/2, gamma4 = 1 - gamma3,
// and per layer the exact solution of
//   dU/dt = gamma1 U - gamma2 D - w F0 gamma3 e^{-tau/mu0} - (gamma1-gamma2) pi B(t)
//   dD/dt = gamma2 U - gamma1 D + w F0 gamma4 e^{-tau/mu0} + (gamma1-gamma2) pi B(t)
// (t the scaled depth in the layer, B linear in t): the homogeneous modes
// A e^{lambda (t - dtau)} (1, Gamma) + B e^{-lambda t} (Gamma, 1), the beam
// particular solution (mu0 moved by 0.1% within 1e-4 of resonance) and the
// thermal one pi B(t) +- pi B' / (gamma1 + gamma2).  No diffuse flux enters
// at the top; the surface is Lambertian with emission (1 - a) pi B_sfc.
// The 2L constants follow from the interface continuity by elimination
// with partial pivoting on each layer's 3 x 2 panel (the carried equation
// and the two continuity equations), keeping x_l = M_l x_{l+1} + m_l for the
// back substitution.  Output row i of ts (3 (L+1) + 1 doubles): U at the
// levels, D at the levels (without the unscaled direct beam, as the solver
// reports it), then the band Planck values B(T_v) of the levels and
// B(T_sfc) (computed first by k_planck_levels, one thread per value, and
// shared with the feature kernel).
// work: 12 doubles per (layer, column), layout [k][layer][column] so that
// the per-column threads of a wavefront touch consecutive addresses: M_l,
// m_l, Gamma_l, E_l, the particular solutions at the layer top and the
// scaled and unscaled depths of the layer top; the back substitution
// overwrites slots 8-9 of layer l with U, D of level l (slots 0-1 of the
// last layer with those of the surface level), and k_ts_pack moves them
// into the rows of ts through LDS.
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrt3 = 1.7320508075688772;

struct TsLayer {
  double gam, E, up0, up1, dn0, dn1, g1, g2, T0, T1;
};

__device__ __forceinline__ TsLayer ts_layer(double dtau, double ssa, double g, double mu0, double f0, double T0,
                                            double bt, double bb, bool thermal) {
  TsLayer r;
  const double w0 = fmin(ssa, 1.0 - 1e-6);
  const double f = g * g;
  const double sc = 1.0 - w0 * f;
  const double dt = sc * dtau;
  const double w = (1.0 - f) * w0 * (1.0 / sc);
  const double gs = g / (1.0 + g);
  const double g1 = kSqrt3 * (2.0 - w * (1.0 + gs)) / 2.0;
  const double g2 = kSqrt3 * w * (1.0 - gs) / 2.0;
  const double g3 = (1.0 - kSqrt3 * gs * mu0) / 2.0;
  const double g4 = 1.0 - g3;
  const double lam = sqrt((g1 - g2) * (g1 + g2));
  r.gam = g2 / (g1 + lam);
  r.E = exp(-lam * dt);
  r.g1 = g1;
  r.g2 = g2;
  r.T0 = T0;
  r.T1 = T0 + dt;
  double m = mu0;
  const double ml = m * lam;
  if (fabs(1.0 - ml * ml) < 1e-4) m *= 0.999;
  const double rm = 1.0 / m;
  const double rdet = 1.0 / (rm * rm - lam * lam);
  const double a = f0 * w * (g3 * (rm - g1) - g2 * g4) * rdet;
  const double b = f0 * w * (-g4 * (g1 + rm) - g2 * g3) * rdet;
  const double e0 = exp(-T0 * rm), e1 = exp(-r.T1 * rm);
  r.up0 = a * e0;
  r.up1 = a * e1;
  r.dn0 = b * e0;
  r.dn1 = b * e1;
  if (thermal) {
    const double s = (bb - bt) / (dt * (g1 + g2));
    r.up0 += bt + s;
    r.up1 += bb + s;
    r.dn0 += bt - s;
    r.dn1 += bb - s;
  }
  return r;
}

// Per-column inputs gathered once into the rows of ts (one thread per
// value): the Planck values of the levels and of the surface (thermal
// batches), and the asymmetry factors g_l = pmom_l[1] (whose moment rows
// are 8*pstride bytes apart, so each is a separate cache line of pmom).
__global__ void k_ts_prep(int64_t n, int L, int P, int thermal, const double* __restrict__ pmom,
                          const double* __restrict__ temper, const double* __restrict__ scal, double* __restrict__ ts) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nb = 2 * L + 2;  // L + 2 Planck values, L asymmetry factors
  if (e >= n * nb) return;
  const int64_t i = e / nb;
  const int v = (int)(e % nb);
  double* row = ts + i * ts_row(L) + 2 * (L + 1);
  if (v < L + 2) {
    if (!thermal) {
      row[v] = 0.0;  // no thermal source
      return;
    }
    const double* sc = scal + i * DISORT_NSCAL;
    const double t = (v <= L) ? temper[i * (L + 1) + v] : sc[DISORT_S_TEMP_BOT];
    row[v] = dsa::planck_band(sc[DISORT_S_WN_LO], sc[DISORT_S_WN_HI], t);
  } else {
    row[v] = pmom[(i * L + (v - L - 2)) * P + 1];
  }
}

__global__ __launch_bounds__(256) void k_twostream(int64_t n, int L, int P, const double* __restrict__ dtau,
                                                   const double* __restrict__ ssa, const double* __restrict__ pmom,
                                                   const double* __restrict__ temper, const double* __restrict__ scal,
                                                   int thermal, double* __restrict__ work, double* __restrict__ ts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* sc = scal + i * DISORT_NSCAL;
  const double mu0 = sc[DISORT_S_BEAM_MU], f0 = sc[DISORT_S_BEAM_FLUX], al = sc[DISORT_S_ALBEDO];
  const bool th = thermal != 0;
  auto W = [&](int k, int l) -> double& { return work[((int64_t)k * L + l) * n + i]; };
  const double* bl = ts + i * ts_row(L) + 2 * (L + 1);  // level Planck values, then the surface's
  const double* gl = bl + L + 2;                          // asymmetry factors
  auto planck_lev = [&](int v) { return kPi * bl[v]; };
  double bt = th ? planck_lev(0) : 0.0;
  double bb = th ? planck_lev(1) : 0.0;
  TsLayer cur = ts_layer(dtau[i * L], ssa[i * L], gl[0], mu0, f0, 0.0, bt, bb, th);
  // Carried equation c x_l = d; at the top D(0) = 0.
  double c0 = cur.gam * cur.E, c1 = 1.0, d = -cur.dn0;
  double Tu = 0.0;  // unscaled depth of the current layer's top
  for (int l = 0; l + 1 < L; ++l) {
    bt = bb;
    bb = th ? planck_lev(l + 2) : 0.0;
    const TsLayer nx = ts_layer(dtau[i * L + l + 1], ssa[i * L + l + 1], gl[l + 1], mu0, f0, cur.T1, bt, bb, th);
    // rows [u0, u1, v0, v1, rhs]: carried, continuity of U, continuity of D
    double R[3][5] = {{c0, c1, 0.0, 0.0, d},
                      {1.0, cur.gam * cur.E, -nx.E, -nx.gam, nx.up0 - cur.up1},
                      {cur.gam, cur.E, -nx.gam * nx.E, -1.0, nx.dn0 - cur.dn1}};
    int p = 0;
    if (fabs(R[1][0]) > fabs(R[p][0])) p = 1;
    if (fabs(R[2][0]) > fabs(R[p][0])) p = 2;
    const int o1 = (p == 0) ? 1 : 0, o2 = (p == 2) ? 1 : 2;
    double A[5], Bq[5], Cr[5];
    const double rp = 1.0 / R[p][0];
    const double fa = R[o1][0] * rp, fb = R[o2][0] * rp;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      A[k] = R[o1][k] - fa * R[p][k];
      Bq[k] = R[o2][k] - fb * R[p][k];
    }
    // second pivot among the two reduced rows (first largest)
    const bool swap = fabs(Bq[1]) > fabs(A[1]);
    double Q[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      Q[k] = swap ? Bq[k] : A[k];
      Cr[k] = swap ? A[k] : Bq[k];
    }
    const double rq = 1.0 / Q[1];
    const double fc = Cr[1] * rq;
#pragma unroll
    for (int k = 2; k < 5; ++k) Cr[k] -= fc * Q[k];
    // x1 = (Q4 - Q2 y0 - Q3 y1) / Q1; x0 = (P4 - P1 x1 - P2 y0 - P3 y1) / P0
    const double M10 = -Q[2] * rq, M11 = -Q[3] * rq, m1 = Q[4] * rq;
    const double* Pr = R[p];
    const double M00 = (-Pr[2] - Pr[1] * M10) * rp, M01 = (-Pr[3] - Pr[1] * M11) * rp;
    const double m0 = (Pr[4] - Pr[1] * m1) * rp;
    W(0, l) = M00;
    W(1, l) = M01;
    W(2, l) = M10;
    W(3, l) = M11;
    W(4, l) = m0;
    W(5, l) = m1;
    W(6, l) = cur.gam;
    W(7, l) = cur.E;
    W(8, l) = cur.up0;
    W(9, l) = cur.dn0;
    W(10, l) = cur.T0;
    W(11, l) = Tu;
    Tu += dtau[i * L + l];
    c0 = Cr[2];
    c1 = Cr[3];
    d = Cr[4];
    cur = nx;
  }
  // Last layer: carried equation and the surface.
  double x0, x1;
  {
    const double b0 = 1.0 - al * cur.gam, b1 = (cur.gam - al) * cur.E;
    double rb = al * mu0 * f0 * exp(-cur.T1 / mu0) - cur.up1 + al * cur.dn1;
    if (th) rb += (1.0 - al) * kPi * bl[L + 1];
    double u0 = c0, u1 = c1, r0 = d, v0 = b0, v1 = b1, r1 = rb;
    if (fabs(v0) > fabs(u0)) {
      double t;
      t = u0, u0 = v0, v0 = t;
      t = u1, u1 = v1, v1 = t;
      t = r0, r0 = r1, r1 = t;
    }
    const double fct = v0 / u0;
    v1 -= fct * u1;
    r1 -= fct * r0;
    x1 = r1 / v1;
    x0 = (r0 - u1 * x1) / u0;
  }
  // Diffuse down as the solver reports it: the scaled total minus the
  // unscaled direct beam, i.e. + mu0 F0 (e^{-T/mu0} - e^{-Tu/mu0}).
  const double u = mu0 * f0;
  auto beam_fix = [&](double T, double Tuv) { return u * (exp(-T / mu0) - exp(-Tuv / mu0)); };
  const double TuL = Tu + dtau[i * L + L - 1];
  W(0, L - 1) = x0 + x1 * cur.gam * cur.E + cur.up1;
  W(1, L - 1) = x0 * cur.gam + x1 * cur.E + cur.dn1 + beam_fix(cur.T1, TuL);
  W(8, L - 1) = x0 * cur.E + x1 * cur.gam + cur.up0;
  W(9, L - 1) = x0 * cur.gam * cur.E + x1 + cur.dn0 + beam_fix(cur.T0, Tu);
  for (int l = L - 2; l >= 0; --l) {
    const double y0 = x0, y1 = x1;
    x0 = W(0, l) * y0 + W(1, l) * y1 + W(4, l);
    x1 = W(2, l) * y0 + W(3, l) * y1 + W(5, l);
    const double gm = W(6, l), E = W(7, l);
    const double uv = x0 * E + x1 * gm + W(8, l);
    const double dv = x0 * gm * E + x1 + W(9, l) + beam_fix(W(10, l), W(11, l));
    W(8, l) = uv;
    W(9, l) = dv;
  }
}

// Two-stream responses to a unit Planck value at each degree of freedom k
// (levels 0..L, then the surface), no beam, one thread per column.  The
// system matrix does not depend on the source: the layer eliminations of
// k_twostream are done once, keeping per layer M_l and the 3 x 3 map T_l
// from the panel's right-hand sides (carried d, U and D continuity) to
// (m_l, carried d'); the surface solve's 2 x 2 map likewise.  Each k then
// only pushes its right-hand side -- non-zero next to level k, where the
// linear thermal source of the two adjacent layers has its particular
// solution pi B(t) +- pi B'/(gamma1 + gamma2) -- through the maps and the
// surface solve: resp[i][0][k] = U at the top (through the running
// products of the M_l, no back substitution), resp[i][1][k] = D at the
// surface.  work: 18 doubles per (layer, column), [slot][layer][column].
__global__ __launch_bounds__(256) void k_twostream_resp(int64_t n, int L, const double* __restrict__ dtau,
                                                        const double* __restrict__ ssa,
                                                        const double* __restrict__ scal,
                                                        const double* __restrict__ ts, double* __restrict__ work,
                                                        double* __restrict__ resp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* sc = scal + i * DISORT_NSCAL;
  const double mu0 = sc[DISORT_S_BEAM_MU], al = sc[DISORT_S_ALBEDO];
  const double* gl = ts + i * ts_row(L) + 3 * (L + 1) + 1;  // asymmetry factors
  auto W = [&](int k, int l) -> double& { return work[((int64_t)k * L + l) * n + i]; };
  // Phase 1: the eliminations with no source.
  TsLayer cur = ts_layer(dtau[i * L], ssa[i * L], gl[0], mu0, 0.0, 0.0, 0.0, 0.0, false);
  double c0 = cur.gam * cur.E, c1 = 1.0;
  for (int l = 0; l < L; ++l) {
    const double dtl = cur.T1 - cur.T0;
    W(13, l) = cur.gam;
    W(14, l) = cur.E;
    W(15, l) = 1.0 / (dtl * (cur.g1 + cur.g2));
    if (l + 1 == L) break;
    const TsLayer nx = ts_layer(dtau[i * L + l + 1], ssa[i * L + l + 1], gl[l + 1], mu0, 0.0, cur.T1, 0.0, 0.0, false);
    const double R[3][4] = {{c0, c1, 0.0, 0.0},
                            {1.0, cur.gam * cur.E, -nx.E, -nx.gam},
                            {cur.gam, cur.E, -nx.gam * nx.E, -1.0}};
    int p = 0;
    if (fabs(R[1][0]) > fabs(R[p][0])) p = 1;
    if (fabs(R[2][0]) > fabs(R[p][0])) p = 2;
    const int o1 = (p == 0) ? 1 : 0, o2 = (p == 2) ? 1 : 2;
    const double rp = 1.0 / R[p][0];
    const double fa = R[o1][0] * rp, fb = R[o2][0] * rp;
    double A[4], Bq[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      A[k] = R[o1][k] - fa * R[p][k];
      Bq[k] = R[o2][k] - fb * R[p][k];
    }
    const bool swap = fabs(Bq[1]) > fabs(A[1]);
    double Q[4], Cr[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      Q[k] = swap ? Bq[k] : A[k];
      Cr[k] = swap ? A[k] : Bq[k];
    }
    const double rq = 1.0 / Q[1];
    const double fc = Cr[1] * rq;
    Cr[2] -= fc * Q[2];
    Cr[3] -= fc * Q[3];
    const double M10 = -Q[2] * rq, M11 = -Q[3] * rq;
    const double M00 = (-R[p][2] - R[p][1] * M10) * rp, M01 = (-R[p][3] - R[p][1] * M11) * rp;
    W(0, l) = M00;
    W(1, l) = M01;
    W(2, l) = M10;
    W(3, l) = M11;
    // T_l: the same row operations on each unit right-hand side.
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double r[3] = {j == 0 ? 1.0 : 0.0, j == 1 ? 1.0 : 0.0, j == 2 ? 1.0 : 0.0};
      const double a4 = r[o1] - fa * r[p], b4 = r[o2] - fb * r[p];
      const double q4 = swap ? b4 : a4;
      const double c4 = (swap ? a4 : b4) - fc * q4;
      const double m1 = q4 * rq;
      const double m0 = (r[p] - R[p][1] * m1) * rp;
      W(4 + j, l) = m0;
      W(7 + j, l) = m1;
      W(10 + j, l) = c4;
    }
    c0 = Cr[2];
    c1 = Cr[3];
    cur = nx;
  }
  // Surface row and the 2 x 2 solve as a map of (d, rb).
  double S00, S01, S10, S11;
  {
    const double b0 = 1.0 - al * cur.gam, b1 = (cur.gam - al) * cur.E;
    const bool sw = fabs(b0) > fabs(c0);
    const double u0 = sw ? b0 : c0, u1 = sw ? b1 : c1, v0 = sw ? c0 : b0, v1 = sw ? c1 : b1;
    const double fct = v0 / u0, v1r = v1 - fct * u1;
    // rows: r0 = (sw ? rb : d), r1 = (sw ? d : rb) - fct r0; x1 = r1 / v1r; x0 = (r0 - u1 x1) / u0
    auto solve = [&](double d, double rb, double& x0, double& x1) {
      const double r0 = sw ? rb : d;
      const double r1 = (sw ? d : rb) - fct * r0;
      x1 = r1 / v1r;
      x0 = (r0 - u1 * x1) / u0;
    };
    solve(1.0, 0.0, S00, S10);
    solve(0.0, 1.0, S01, S11);
  }
  // U at the top is [E_0, Gamma_0] x_0 + up_0 with x_0 = sum_l M_0..M_{l-1} m_l
  // + M_0..M_{L-2} x_{L-1}: keep the row vectors w_l = [E_0, Gamma_0]
  // M_0..M_{l-1} (slots 16, 17) so that each right-hand side needs no back
  // substitution.
  {
    double w0 = W(14, 0), w1 = W(13, 0);
    for (int l = 0; l < L; ++l) {
      W(16, l) = w0;
      W(17, l) = w1;
      if (l + 1 == L) break;
      const double n0 = w0 * W(0, l) + w1 * W(2, l), n1 = w0 * W(1, l) + w1 * W(3, l);
      w0 = n0;
      w1 = n1;
    }
  }
  // Phase 2: the right-hand sides of G degrees of freedom at a time, from
  // the first interface where any of them is non-zero (two above level k).
  constexpr int G = 4;
  const int D = L + 2;
  for (int k0 = 0; k0 < D; k0 += G) {
    // particular solution of layer l at its top (0) and bottom (1) for B = e_k
    auto part = [&](int k, int l, double& up0, double& up1, double& dn0, double& dn1) {
      const double bt = (l == k) ? kPi : 0.0, bb = (l + 1 == k) ? kPi : 0.0;
      const double s = (bb - bt) * W(15, l);
      up0 = bt + s;
      up1 = bb + s;
      dn0 = bt - s;
      dn1 = bb - s;
    };
    const int l0 = (k0 >= 2) ? k0 - 2 : 0;
    double d[G], u1a[G], d1a[G], acc[G], up00[G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
      double u0a, d0a;
      part(k0 + q, l0, u0a, u1a[q], d0a, d1a[q]);
      d[q] = (l0 == 0) ? -d0a : 0.0;
      up00[q] = (l0 == 0) ? u0a : 0.0;
      acc[q] = 0.0;
    }
    for (int l = l0; l + 1 < L; ++l) {
      double T[9];
#pragma unroll
      for (int t = 0; t < 9; ++t) T[t] = W(4 + t, l);
      const double wl0 = W(16, l), wl1 = W(17, l);
#pragma unroll
      for (int q = 0; q < G; ++q) {
        double u0b, u1b, d0b, d1b;
        part(k0 + q, l + 1, u0b, u1b, d0b, d1b);
        const double r1 = u0b - u1a[q], r2 = d0b - d1a[q];
        const double m0 = T[0] * d[q] + T[1] * r1 + T[2] * r2;
        const double m1 = T[3] * d[q] + T[4] * r1 + T[5] * r2;
        d[q] = T[6] * d[q] + T[7] * r1 + T[8] * r2;
        acc[q] += wl0 * m0 + wl1 * m1;
        u1a[q] = u1b;
        d1a[q] = d1b;
      }
    }
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int k = k0 + q;
      if (k >= D) break;
      const double rb = -u1a[q] + al * d1a[q] + ((k == L + 1) ? (1.0 - al) * kPi : 0.0);
      const double x0 = S00 * d[q] + S01 * rb, x1 = S10 * d[q] + S11 * rb;
      resp[i * 2 * D + D + k] = x0 * W(13, L - 1) + x1 * W(14, L - 1) + d1a[q];
      resp[i * 2 * D + k] = acc[q] + W(16, L - 1) * x0 + W(17, L - 1) * x1 + up00[q];
    }
  }
}

// Rows of ts from the per-level results the back substitution left in the
// work array: 32 columns per workgroup, staged through LDS so that the
// reads ([k][level][column]) and the row writes are both coalesced.
__global__ __launch_bounds__(256) void k_ts_pack(int64_t n, int L, const double* __restrict__ work,
                                                 double* __restrict__ ts) {
  __shared__ double tile[32 * 2 * 64];
  const int nlev = L + 1, K = 2 * nlev;
  const int64_t i0 = (int64_t)blockIdx.x * 32;
  for (int e = threadIdx.x; e < K * 32; e += blockDim.x) {
    const int k = e / 32, c = e % 32;
    const int64_t i = i0 + c;
    if (i >= n) continue;
    const int q = k / nlev, v = k % nlev;  // q 0: U, 1: D
    const int slot = (v < L) ? 8 + q : q, l = (v < L) ? v : L - 1;
    tile[c * K + k] = work[((int64_t)slot * L + l) * n + i];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < K * 32; e += blockDim.x) {
    const int c = e / K, k = e % K;
    const int64_t i = i0 + c;
    if (i < n) ts[i * ts_row(L) + k] = tile[c * K + k];
  }
}

// ---------------------------------------------------------------------------
// Jacobian surrogate: the temperature Jacobians of the top-of-atmosphere
// upward flux and the surface downward flux, dF/dT_k = R_k dB/dT(T_k) for the
// D = nlyr + 2 degrees of freedom k (level temperatures, then the surface).
// The responses R_k = dF/dB_k depend on the optical properties only; the
// network learns R / pi and dB/dT is evaluated exactly.
// Emission features of column i (3 nlyr + 1 + 6 (nlyr+1), zero padding to w0),
// with mubar = 0.6 the diffusivity cosine:
//   [l] log10(delta_tau_l), [nlyr + l] single_scat_albedo_l, [2nlyr + l] g_l,
//   [3nlyr] surface albedo, then six blocks over the levels v = 0..nlyr:
//   log10(tau_v + 1e-3), log10(a_v + 1e-3), exp(-tau_v / mubar),
//   exp(-a_v / mubar), exp(-(tau_L - tau_v) / mubar), exp(-(a_L - a_v) / mubar)
// with tau_v the optical depth and a_v = sum_{l<v} (1 - omega_l) dtau_l the
// absorption optical depth of level v (L = nlyr: the surface level).
// Targets (2D, then zero padding to wL):
//   [k]      R_k / pi of flux_up at the top level
//   [D + k]  R_k / pi of flux_down at the surface level
// ---------------------------------------------------------------------------
constexpr double kMuBar = 0.6;

__device__ __forceinline__ double dbdt_dof(const double* temper, const double* scal, int64_t i, int L, int k) {
  const double* sc = scal + i * DISORT_NSCAL;
  const double t = (k <= L) ? temper[i * (L + 1) + k] : sc[DISORT_S_TEMP_BOT];
  return dsa::planck_band_deriv(sc[DISORT_S_WN_LO], sc[DISORT_S_WN_HI], t);
}

// Offset of the Jacobian row of target row r (0: flux_up at level 0, 1:
// flux_down at level L) within column i's [6][L+1][D] block.
__device__ __forceinline__ int64_t jac_row(int64_t i, int L, int r) {
  const int D = L + 2;
  const int64_t per = (int64_t)(L + 1) * D;
  return i * 6 * per + (r == 0 ? 0 : per + (int64_t)L * D);
}

// One wavefront per column (four per workgroup), as k_features.
__global__ __launch_bounds__(256) void k_features_emission(int64_t n, int64_t n_pad, int L, int P,
                                                           const double* dtau, const double* ssa,
                                                           const double* pmom, const double* scal,
                                                           const float* shift, const float* scale, const float* lo,
                                                           const float* hi, int w0, __bf16* x, int32_t* ood,
                                                           const double* resp) {
  __shared__ double lev[4][2][64];  // [column][tau, absorption tau][level]
  __shared__ double lay[4][2][64];  // [column][dtau, (1 - omega) dtau][layer]
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + wv;
  if (i >= n_pad) return;
  __bf16* xr = x + i * w0;
  if (i >= n) {
    for (int k = lane; k < w0; k += 64) xr[k] = (__bf16)0.0f;
    return;
  }
  const int nlev = L + 1, base = 3 * L + 1, nf0 = base + 6 * nlev, D = L + 2;
  // With the two-stream responses: their 2D values / pi follow.
  const bool with_r = resp && w0 >= nf0 + 2 * D;
  const int nf = nf0 + (with_r ? 2 * D : 0);
  double (*lv)[64] = lev[wv];
  double (*ly)[64] = lay[wv];
  if (lane < L) {
    const double dt = dtau[i * L + lane];
    ly[0][lane] = dt;
    ly[1][lane] = (1.0 - ssa[i * L + lane]) * dt;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane < nlev) {
    double tau = 0.0, tab = 0.0;
    for (int l = 0; l < lane; ++l) {
      tau += ly[0][l];
      tab += ly[1][l];
    }
    lv[0][lane] = tau;
    lv[1][lane] = tab;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const double tauL = lv[0][L], tabL = lv[1][L];
  int bad = 0;
  for (int k = lane; k < w0; k += 64) {
    double f = 0.0;
    if (k < L) f = flog10(dtau[i * L + k]);
    else if (k < 2 * L) f = ssa[i * L + (k - L)];
    else if (k < 3 * L) f = pmom[(i * L + (k - 2 * L)) * P + 1];
    else if (k == 3 * L) f = scal[i * DISORT_NSCAL + DISORT_S_ALBEDO];
    else if (k >= nf0 && k < nf) f = resp[i * 2 * D + (k - nf0)] / kPi;
    else if (k < nf) {
      const int blk = (k - base) / nlev, v = (k - base) % nlev;
      const double tau = lv[0][v], tab = lv[1][v];
      f = (blk == 0)   ? flog10(tau + 1e-3)
          : (blk == 1) ? flog10(tab + 1e-3)
          : (blk == 2) ? fexp(-tau / kMuBar)
          : (blk == 3) ? fexp(-tab / kMuBar)
          : (blk == 4) ? fexp(-(tauL - tau) / kMuBar)
                       : fexp(-(tabL - tab) / kMuBar);
    }
    if (k < nf) {
      const float ff = (float)f;
      bad |= !(ff >= lo[k] && ff <= hi[k]);
      xr[k] = (__bf16)((ff - shift[k]) * scale[k]);
    } else {
      xr[k] = (__bf16)0.0f;
    }
  }
  const unsigned long long any_bad = __ballot(bad);
  if (ood && lane == 0) ood[i] = any_bad ? 1 : 0;
}

__global__ void k_jac_targets(int64_t n, int64_t n_pad, int L, const double* jac, const double* temper,
                              const double* scal, const float* shift, const float* scale, int wL, float* t,
                              const double* resp) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_pad * wL) return;
  const int64_t i = e / wL;
  const int k = (int)(e % wL), D = L + 2;
  float v = 0.f;
  if (i < n && k < 2 * D) {
    const int r = k / D, d = k % D;
    double rr = jac[jac_row(i, L, r) + d] / (kPi * dbdt_dof(temper, scal, i, L, d));
    if (resp) rr -= resp[i * 2 * D + k] / kPi;  // residual over the two-stream response
    v = ((float)rr - shift[k]) * scale[k];
  }
  t[e] = v;
}

// Network outputs -> Jacobian rows jrows [n][2][D] (mean of the members) and
// the members' spread max_k |R_1 - R_2| / pi per column (lane = degree of freedom).
__global__ __launch_bounds__(256) void k_jac_decode(int64_t n, int L, int wL, const float* y1, const float* y2,
                                                    const float* shift, const float* scale, const double* temper,
                                                    const double* scal, double* jrows, float* spread,
                                                    const double* resp) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const int D = L + 2;
  double worst = 0.0;
  if (lane < D) {
    const double db = kPi * dbdt_dof(temper, scal, i, L, lane);
    for (int r = 0; r < 2; ++r) {
      const int k = r * D + lane;
      const double base = resp ? resp[i * 2 * D + k] / kPi : 0.0;  // two-stream response (R / pi)
      const double a1 = (double)y1[i * wL + k] / scale[k] + shift[k] + base;
      double a = a1;
      if (y2) {
        const double a2 = (double)y2[i * wL + k] / scale[k] + shift[k] + base;
        a = 0.5 * (a1 + a2);
        worst = fmax(worst, fabs(a1 - a2));
        if (!(a1 - a2 == a1 - a2)) worst = __builtin_inf();
      }
      jrows[(i * 2 + r) * D + lane] = a * db;
    }
  }
  for (int off = 32; off >= 1; off >>= 1) worst = fmax(worst, __shfl_xor(worst, off));
  if (lane == 0) spread[i] = (float)worst;
}

// Online error monitor of the Jacobian gate: sampled columns that pass the
// gate, error max over both rows and every k of |dF/dT_k - ref| / (pi dB/dT(T_k)).
__global__ void k_jac_monitor(int64_t m, const int32_t* idx, const int32_t* flags, int L, const double* jac_sub,
                              const double* jrows, const double* temper, const double* scal, double tol,
                              int32_t* counters) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool smp = false, bad = false;
  if (k < m) {
    const int64_t i = idx[k];
    smp = (flags[i] & 6) == 6;
    if (smp) {
      const int D = L + 2;
      double worst = 0.0;
      for (int d = 0; d < D; ++d) {
        const double den = kPi * dbdt_dof(temper, scal, i, L, d);
        for (int r = 0; r < 2; ++r)
          worst = fmax(worst, fabs(jrows[(i * 2 + r) * D + d] - jac_sub[jac_row(k, L, r) + d]) / den);
      }
      bad = !(worst <= tol);
    }
  }
  const unsigned long long bs = __ballot(smp), bb = __ballot(bad);
  if ((threadIdx.x & 63) == 0) {
    if (bs) atomicAdd(&counters[0], __popcll(bs));
    if (bb) atomicAdd(&counters[1], __popcll(bb));
  }
}

// jrows of column idx[k] <- the two rows of sub-batch column k's Jacobian.
__global__ void k_jac_scatter(int64_t m, const int32_t* idx, int L, const double* jac_sub, double* jrows) {
  const int D = L + 2;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m * 2 * D) return;
  const int64_t k = e / (2 * D);
  const int r = (int)((e / D) % 2), d = (int)(e % D);
  jrows[((int64_t)idx[k] * 2 + r) * D + d] = jac_sub[jac_row(k, L, r) + d];
}

}  // namespace

extern "C" {

#define SURR_LAUNCH_CHECK()                                          \
  do {                                                               \
    if (hipGetLastError() != hipSuccess) return DISORT_E_HIP;        \
  } while (0)

int disort_surrogate_features(disort_ctx* ctx, int64_t n, int64_t n_pad, int nlyr, int pstride,
                              const disort_flux_batch* b, const float* shift, const float* scale, const float* lo,
                              const float* hi, int w0, uint16_t* x, int32_t* ood, const double* ts) {
  if (!ctx || !b || n < 0 || n_pad < n || nlyr <= 0 || pstride < 2 || w0 < 4 * nlyr + 2) return DISORT_E_INVALID;
  if (!b->temper) return DISORT_E_INVALID;
  if (nlyr + 2 > 64) return DISORT_E_INVALID;  // one lane per level
  hipLaunchKernelGGL(k_features, dim3(blocks(n_pad, 4)), dim3(256), 0, (hipStream_t)disort_ctx_stream(ctx), n, n_pad, nlyr,
                     pstride, b->dtau, b->ssa, b->pmom, b->temper, b->scal, shift, scale, lo, hi, w0, (__bf16*)x, ood,
                     ts);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_surrogate_targets(disort_ctx* ctx, int64_t n, int64_t n_pad, int nlev, const disort_flux_batch* b,
                             const float* shift, const float* scale, int wL, float* t, const double* ts) {
  if (!ctx || !b || !b->out || n < 0 || n_pad < n || wL < 3 * nlev) return DISORT_E_INVALID;
  hipLaunchKernelGGL(k_targets, dim3(blocks(n_pad * wL)), dim3(256), 0, (hipStream_t)disort_ctx_stream(ctx), n, n_pad,
                     nlev, b->out, b->total_levels, b->scal, shift, scale, wL, t, ts);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_surrogate_decode(disort_ctx* ctx, int64_t n, int nlyr, int wL, const float* const* ys, int n_members,
                            const float* shift, const float* scale, const disort_flux_batch* b, double c_local,
                            double c_floor, double* flux, float* spread, const double* ts) {
  if (!ctx || !b || !ys || !flux || !spread || wL < 3 * (nlyr + 1) || !(c_floor > 0.0)) return DISORT_E_INVALID;
  if (n_members < 1 || n_members > kMaxMembers) return DISORT_E_INVALID;
  if (nlyr + 1 > 64) return DISORT_E_INVALID;  // one lane per level
  Members ms{};
  ms.k = n_members;
  for (int m = 0; m < n_members; ++m) {
    if (!ys[m]) return DISORT_E_INVALID;
    ms.y[m] = ys[m];
  }
  hipLaunchKernelGGL(k_decode, dim3(blocks(n, 4)), dim3(256), 0, (hipStream_t)disort_ctx_stream(ctx), n, nlyr, wL, ms,
                     shift, scale, b->dtau, b->scal, c_local, c_floor, flux, spread, ts);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_surrogate_gather(disort_ctx* ctx, int64_t m, const int32_t* idx, int nlyr, int pstride,
                            const disort_flux_batch* src, disort_flux_batch* dst) {
  if (!ctx || !idx || !src || !dst || m < 0) return DISORT_E_INVALID;
  if (m == 0) return DISORT_OK;
  hipStream_t st = (hipStream_t)disort_ctx_stream(ctx);
  auto rows = [&](int64_t elems, const void* s, void* d) {
    hipLaunchKernelGGL(k_rows, dim3(blocks(m * elems)), dim3(256), 0, st, m, idx, elems, 8, s, d);
  };
  rows(nlyr, src->dtau, (void*)dst->dtau);
  rows(nlyr, src->ssa, (void*)dst->ssa);
  rows((int64_t)nlyr * pstride, src->pmom, (void*)dst->pmom);
  if (src->temper) rows(nlyr + 1, src->temper, (void*)dst->temper);
  rows(DISORT_NSCAL, src->scal, (void*)dst->scal);
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

int disort_surrogate_route(disort_ctx* ctx, int64_t n, const float* spread, const int32_t* ood, float tau,
                           double sample_frac, uint64_t seed, int route_all, int32_t* flags, int32_t* idx,
                           int32_t* count) {
  if (!ctx || !spread || !flags || !idx || !count || n < 0) return DISORT_E_INVALID;
  hipStream_t st = (hipStream_t)disort_ctx_stream(ctx);
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, count, 1);
  if (n > 0)
    hipLaunchKernelGGL(k_route, dim3(blocks(n)), dim3(256), 0, st, n, spread, ood, tau, sample_frac, seed, route_all,
                       flags, idx, count);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_surrogate_pad_index(disort_ctx* ctx, int64_t m, int64_t mpad, int32_t* idx) {
  if (!ctx || !idx || m <= 0 || mpad < m) return DISORT_E_INVALID;
  if (mpad > m)
    hipLaunchKernelGGL(k_pad_idx, dim3(blocks(mpad - m)), dim3(256), 0, (hipStream_t)disort_ctx_stream(ctx), m, mpad, idx);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_surrogate_monitor(disort_ctx* ctx, int64_t m, const int32_t* idx, const int32_t* flags, int nlev,
                             const double* out, int64_t total_levels, const double* pred, const disort_flux_batch* b,
                             double c_local, double c_floor, double tol, int32_t* counters) {
  if (!ctx || !idx || !flags || !out || !pred || !b || !counters || m < 0 || !(c_floor > 0.0)) return DISORT_E_INVALID;
  hipStream_t st = (hipStream_t)disort_ctx_stream(ctx);
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, counters, 2);
  if (m > 0)
    hipLaunchKernelGGL(k_monitor, dim3(blocks(m)), dim3(256), 0, st, m, idx, flags, nlev, out, total_levels, pred, b->scal,
                       c_local, c_floor, tol, counters);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_surrogate_twostream(disort_ctx* ctx, int64_t n, int nlyr, int pstride, const disort_flux_batch* b,
                               double* work, double* ts) {
  if (!ctx || !b || !work || !ts || n < 0 || nlyr <= 0 || nlyr + 1 > 64 || pstride < 2) return DISORT_E_INVALID;
  const bool thermal = (b->flags & DISORT_F_THERMAL) != 0;
  if (thermal && !b->temper) return DISORT_E_INVALID;
  if (n == 0) return DISORT_OK;
  hipStream_t st = (hipStream_t)disort_ctx_stream(ctx);
  hipLaunchKernelGGL(k_ts_prep, dim3(blocks(n * (2 * nlyr + 2))), dim3(256), 0, st, n, nlyr, pstride, thermal ? 1 : 0,
                     b->pmom, b->temper, b->scal, ts);
  hipLaunchKernelGGL(k_twostream, dim3(blocks(n)), dim3(256), 0, st, n, nlyr, pstride, b->dtau, b->ssa, b->pmom,
                     b->temper, b->scal, thermal ? 1 : 0, work, ts);
  hipLaunchKernelGGL(k_ts_pack, dim3(blocks(n, 32)), dim3(256), 0, st, n, nlyr, work, ts);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_jac_twostream(disort_ctx* ctx, int64_t n, int nlyr, int pstride, const disort_flux_batch* b, double* work,
                         int64_t work_cols, double* ts, double* resp) {
  if (!ctx || !b || !work || !ts || !resp || n < 0 || nlyr <= 0 || nlyr + 2 > 64 || pstride < 2 || work_cols < 1)
    return DISORT_E_INVALID;
  hipStream_t st = (hipStream_t)disort_ctx_stream(ctx);
  const int L = nlyr, D = nlyr + 2, R = ts_row(nlyr);
  for (int64_t c0 = 0; c0 < n; c0 += work_cols) {
    const int64_t m = std::min(work_cols, n - c0);
    const double* dtau = b->dtau + c0 * L;
    const double* ssa = b->ssa + c0 * L;
    const double* pmom = b->pmom + c0 * L * pstride;
    const double* scal = b->scal + c0 * DISORT_NSCAL;
    double* tsc = ts + c0 * R;
    // the columns' g (no Planck values: the sources are unit vectors)
    hipLaunchKernelGGL(k_ts_prep, dim3(blocks(m * (2 * L + 2))), dim3(256), 0, st, m, L, pstride, 0, pmom, nullptr,
                       scal, tsc);
    hipLaunchKernelGGL(k_twostream_resp, dim3(blocks(m)), dim3(256), 0, st, m, L, dtau, ssa, scal, tsc, work,
                       resp + c0 * 2 * D);
  }
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_jac_surrogate_features(disort_ctx* ctx, int64_t n, int64_t n_pad, int nlyr, int pstride,
                                  const disort_flux_batch* b, const float* shift, const float* scale, const float* lo,
                                  const float* hi, int w0, uint16_t* x, int32_t* ood, const double* resp) {
  if (!ctx || !b || n < 0 || n_pad < n || nlyr <= 0 || pstride < 2 || w0 < 3 * nlyr + 1 + 6 * (nlyr + 1))
    return DISORT_E_INVALID;
  if (nlyr + 1 > 64) return DISORT_E_INVALID;  // one lane per level
  hipLaunchKernelGGL(k_features_emission, dim3(blocks(n_pad, 4)), dim3(256), 0, (hipStream_t)disort_ctx_stream(ctx), n,
                     n_pad, nlyr, pstride, b->dtau, b->ssa, b->pmom, b->scal, shift, scale, lo, hi, w0, (__bf16*)x, ood,
                     resp);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_jac_surrogate_targets(disort_ctx* ctx, int64_t n, int64_t n_pad, int nlyr, const double* jac,
                                 const disort_flux_batch* b, const float* shift, const float* scale, int wL,
                                 float* t, const double* resp) {
  if (!ctx || !b || !jac || !b->temper || n < 0 || n_pad < n || wL < 2 * (nlyr + 2)) return DISORT_E_INVALID;
  hipLaunchKernelGGL(k_jac_targets, dim3(blocks(n_pad * wL)), dim3(256), 0, (hipStream_t)disort_ctx_stream(ctx), n,
                     n_pad, nlyr, jac, b->temper, b->scal, shift, scale, wL, t, resp);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_jac_surrogate_decode(disort_ctx* ctx, int64_t n, int nlyr, int wL, const float* y1, const float* y2,
                                const float* shift, const float* scale, const disort_flux_batch* b, double* jrows,
                                float* spread, const double* resp) {
  if (!ctx || !b || !b->temper || !y1 || !jrows || !spread || wL < 2 * (nlyr + 2)) return DISORT_E_INVALID;
  if (nlyr + 2 > 64) return DISORT_E_INVALID;  // one lane per degree of freedom
  hipLaunchKernelGGL(k_jac_decode, dim3(blocks(n, 4)), dim3(256), 0, (hipStream_t)disort_ctx_stream(ctx), n, nlyr, wL,
                     y1, y2, shift, scale, b->temper, b->scal, jrows, spread, resp);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_jac_surrogate_monitor(disort_ctx* ctx, int64_t m, const int32_t* idx, const int32_t* flags, int nlyr,
                                 const double* jac_sub, const double* jrows, const disort_flux_batch* b, double tol,
                                 int32_t* counters) {
  if (!ctx || !idx || !flags || !jac_sub || !jrows || !b || !b->temper || !counters || m < 0) return DISORT_E_INVALID;
  hipStream_t st = (hipStream_t)disort_ctx_stream(ctx);
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, counters, 2);
  if (m > 0)
    hipLaunchKernelGGL(k_jac_monitor, dim3(blocks(m)), dim3(256), 0, st, m, idx, flags, nlyr, jac_sub, jrows, b->temper,
                       b->scal, tol, counters);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_jac_surrogate_scatter(disort_ctx* ctx, int64_t m, const int32_t* idx, int nlyr, const double* jac_sub,
                                 double* jrows) {
  if (!ctx || !idx || !jac_sub || !jrows || m < 0) return DISORT_E_INVALID;
  if (m == 0) return DISORT_OK;
  hipLaunchKernelGGL(k_jac_scatter, dim3(blocks(m * 2 * (nlyr + 2))), dim3(256), 0, (hipStream_t)disort_ctx_stream(ctx),
                     m, idx, nlyr, jac_sub, jrows);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

int disort_rows_gather(disort_ctx* ctx, int64_t m, const int32_t* idx, int64_t row_elems, int elem_bytes,
                       const void* src, void* dst) {
  if (!ctx || !idx || !src || !dst || m < 0 || (elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8))
    return DISORT_E_INVALID;
  if (m == 0) return DISORT_OK;
  hipLaunchKernelGGL(k_rows, dim3(blocks(m * row_elems)), dim3(256), 0, (hipStream_t)disort_ctx_stream(ctx), m, idx,
                     row_elems, elem_bytes, src, dst);
  SURR_LAUNCH_CHECK();
  return DISORT_OK;
}

}  // extern "C"
