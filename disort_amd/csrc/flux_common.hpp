// Device code shared by the flux kernels of every stream count
// (flux_kernels.hip for N <= 16, flux_wide.hip for N = 32, 64):
// constants, the band-integrated Planck function, lane-broadcast helpers,
// the per-instance set-up kernel k_column, the eigenvector accessor g_entry
// and the flux assembly kernel k_flux.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "../../include/disort_amd.h"
#include "flux_args.hpp"
#include "flux_layout.hpp"
#include "planck.hpp"

namespace dsa {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kEps = 2.220446049250313e-16;     // DBL_EPSILON
constexpr double kLogDblMax = 709.782712893384;    // log(DBL_MAX)

// Broadcast lane `src` of each row of 16 lanes to the whole row (DPP row_newbcast).
template <int SRC>
__device__ __forceinline__ double row_bcast(double x) {
  static_assert(SRC >= 0 && SRC < 16, "row lane");
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), 0x150 + SRC, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), 0x150 + SRC, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// x of the lane given by DPP control CTRL (a pattern within rows of 16 lanes).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// x of lane (this lane xor M), M < 32 (ds_swizzle bit mode: and 0x1f, xor M).
template <int M>
__device__ __forceinline__ double swz_xor(double x) {
  static_assert(M > 0 && M < 32, "swizzle within 32 lanes");
  const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(x), (M << 10) | 0x1F);
  const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(x), (M << 10) | 0x1F);
  return __hiloint2double(hi, lo);
}

// Sum of x over aligned groups of L lanes (L a power of two <= 32); every
// lane of the group gets the same value (the butterfly adds commute).
template <int L>
__device__ __forceinline__ double group_sum(double x) {
  if constexpr (L > 1) x += swz_xor<1>(x);
  if constexpr (L > 2) x += swz_xor<2>(x);
  if constexpr (L > 4) x += swz_xor<4>(x);
  if constexpr (L > 8) x += swz_xor<8>(x);
  if constexpr (L > 16) x += swz_xor<16>(x);
  return x;
}

// Maximum of x over a 16-lane DPP row (every lane gets it): rotations by 8,
// 4, 2, 1 cover the row.  fmax ignores NaN.
__device__ __forceinline__ double row_max(double x) {
  x = fmax(x, dpp_f64<0x128>(x));  // row_ror:8
  x = fmax(x, dpp_f64<0x124>(x));  // row_ror:4
  x = fmax(x, dpp_f64<0x122>(x));  // row_ror:2
  x = fmax(x, dpp_f64<0x121>(x));  // row_ror:1
  return x;
}

// Maximum of x over the aligned 32-lane half of the wavefront, given x
// uniform within each 16-lane row (a row maximum).  v_permlane16_swap with
// both operands x returns the even rows of x in its first result and the odd
// rows in its second, in every lane of each row pair, without the LDS path.
__device__ __forceinline__ double row_pair_max(double x) {
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(x), (unsigned)__double2loint(x), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(x), false, false);
  return fmax(__hiloint2double((int)hi[0], (int)lo[0]), __hiloint2double((int)hi[1], (int)lo[1]));
}

// LDS ordering inside one wavefront: LDS instructions of a wavefront execute
// in order, so only the compiler must not move loads and stores across.
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Compile-time loop helper: f(integral_constant<int, I>) for I in [B, E).
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

__host__ __device__ constexpr int tri(int i, int j) { return i * (i + 1) / 2 + j; }  // i >= j

// Slant optical depth of the pseudo-spherical beam from the top down to
// fraction tau_frac above level lc of layer lc (1-based), through layers
// of thickness dtau between level altitudes zd (top first) on a sphere of
// radius r (src/FluxSolver.cpp:1503-1566).
__device__ double chapman(int lc, double tau_frac, const double* zd, const double* dtau, int nlyr, double r,
                          double mu0) {
  const double deg = kPi / 180.0;
  const double zenang = acos(mu0) * 180.0 / kPi;
  const double zenrad = zenang * deg;
  const double xp = r + zd[lc] + (zd[lc - 1] - zd[lc]) * tau_frac;
  const double xpsinz = xp * sin(zenrad);
  if (zenang > 90.0 && xpsinz < r) return 1.0e+20;
  int id = lc;
  if (zenang > 90.0)
    for (int j = lc; j <= nlyr; ++j)
      if (xpsinz < (zd[j - 1] + r) && xpsinz >= (zd[j] + r)) id = j;
  double sum = 0.0;
  for (int j = 1; j <= id; ++j) {
    double fact = 1.0, fact2 = 1.0;
    if (j > lc) fact = 2.0;
    else if (j == lc && lc == id && zenang > 90.0) fact2 = -1.0;
    const double rj = r + zd[j - 1];
    double rjp1 = r + zd[j];
    if (j == lc && id == lc) rjp1 = xp;
    const double dhj = zd[j - 1] - zd[j];
    double dsj;
    if (id > lc && j == id) dsj = sqrt(rj * rj - xpsinz * xpsinz);
    else dsj = sqrt(rj * rj - xpsinz * xpsinz) - fact2 * sqrt(rjp1 * rjp1 - xpsinz * xpsinz);
    sum += dtau[j - 1] * fact * dsj / dhj;
  }
  if (id > lc) {
    const double dhj = zd[lc - 1] - zd[lc];
    const double dsj = sqrt(xp * xp - xpsinz * xpsinz) - sqrt((zd[lc] + r) * (zd[lc] + r) - xpsinz * xpsinz);
    sum += dtau[lc - 1] * dsj / dhj;
  }
  return sum;
}

// ===========================================================================
// k_column: per-instance set-up (one wavefront per instance).
// ===========================================================================
template <int N>
__global__ __launch_bounds__(64) void k_column(FluxArgs a) {
  const int bk = blockIdx.x;
  if (bk >= a.nb) return;
  const int lane = threadIdx.x;
  const int i = a.b_inst[bk];
  const int L = a.nlyr[i];
  const int P = a.pstride[i];
  const int64_t lo = a.lay_off[i], vo = lo + i;
  const double* pm = a.pmom + a.mom_off[i];
  const double* sc = a.scal + (int64_t)i * DISORT_NSCAL;
  const bool thermal = (a.flags & DISORT_F_THERMAL) != 0;
  const bool want_dmp = (a.flags & DISORT_F_DELTA_M_PLUS) != 0;
  const int64_t TL = a.total_layers, TV = a.total_levels;

  // src/FluxSolver.cpp:56-81: delta-M when any pmom[N] != 0; Delta-M+ only if
  // every layer has pmom[N] >= 1e-4 and pmom[N+1] >= 0.7 pmom[N].
  int any_dm = 0, dmp_bad = 0;
  for (int lc = lane; lc < L; lc += 64) {
    const double pN = pm[(int64_t)lc * P + N];
    any_dm |= (pN != 0.0);
    if (want_dmp) {
      const double pN1 = pm[(int64_t)lc * P + N + 1];
      dmp_bad |= (pN < 1e-4 || pN1 < 0.7 * pN);
    }
  }
  const bool dmp = want_dmp && !__any(dmp_bad);
  const bool deltam = !dmp && __any(any_dm);

  // Per-layer scaling (src/FluxSolver.cpp:333-389).
  for (int lc = lane; lc < L; lc += 64) {
    const int64_t g = lo + lc;
    const double w = a.ssa[g], dt = a.dtau[g];
    double f = 0.0, sig2 = -1.0, ws = w, dts = dt;
    if (deltam || dmp) {
      const double pN = pm[(int64_t)lc * P + N];
      sig2 = 0.0;
      f = pN;
      if (dmp && pm[(int64_t)lc * P + N - 1] != pN) {
        const double pN1 = pm[(int64_t)lc * P + N + 1];
        if (pN != pN1) {
          const double M = (double)N;
          sig2 = ((M + 1.0) * (M + 1.0) - M * M) / (log(pN * pN) - log(pN1 * pN1));
          f = exp(M * M / (2.0 * sig2)) * pN;
        }
      }
      ws = w * (1.0 - f) / (1.0 - f * w);
      dts = (1.0 - f * w) * dt;
    }
    a.lay_d[LD_DTAU_S * TL + g] = dts;
    a.lay_d[LD_SSA_S * TL + g] = ws;
    a.lay_d[LD_F * TL + g] = f;
    a.lay_d[LD_SIG2 * TL + g] = sig2;
  }
  __syncthreads();

  const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU];
  const bool spherical = (a.flags & DISORT_F_SPHERICAL) != 0 && F0 > 0.0;
  __shared__ int s_ncut;
  __shared__ double s_abst;
  if (lane == 0) {
    // Prefix sums and absorption cut-off (:311-331, :428-435).
    double taus = 0.0, tauc = 0.0, abst = 0.0;
    int ncut = L;
    a.lev_d[VD_TAU_S * TV + vo] = 0.0;
    a.lev_d[VD_TAUC * TV + vo] = 0.0;
    for (int lc = 0; lc < L; ++lc) {
      const int64_t g = lo + lc;
      if (abst < 10.0) ncut = lc + 1;
      abst += (1.0 - a.ssa[g]) * a.dtau[g];
      a.lay_d[LD_TAU_TOP * TL + g] = taus;
      taus += a.lay_d[LD_DTAU_S * TL + g];
      tauc += a.dtau[g];
      a.lev_d[VD_TAU_S * TV + vo + lc + 1] = taus;
      a.lev_d[VD_TAUC * TV + vo + lc + 1] = tauc;
    }
    s_ncut = ncut;
    s_abst = abst;
  }
  __syncthreads();
  const int ncut = s_ncut;
  // Beam path factor per layer (:392-421): mu0, or with the spherical beam
  // the Chapman factor at the layer middle and the slant depth at the layer
  // bottom, filled for lc < ncut only (zero below, as in a freshly
  // constructed reference solver).
  for (int lc = lane; lc < L; lc += 64) {
    const int64_t g = lo + lc;
    double ch = mu0, cht = 0.0;
    if (spherical) {
      ch = 0.0;
      if (lc < ncut) {
        const double* zd = a.alt + vo;
        const double* dts = a.lay_d + LD_DTAU_S * TL + lo;
        const double rad = sc[DISORT_S_RADIUS];
        const double taup_mid = a.lay_d[LD_TAU_TOP * TL + g] + dts[lc] / 2.0;
        ch = taup_mid / chapman(lc + 1, 0.5, zd, dts, L, rad, mu0);
        cht = chapman(lc + 1, 0.0, zd, dts, L, rad, mu0);
      }
    }
    a.lay_d[LD_CH * TL + g] = ch;
    a.lay_d[LD_CHTAU * TL + g] = cht;
  }
  __syncthreads();
  if (lane == 0) {
    // Transmission as seen by the boundary conditions: exp(-tau'/mu0) (or
    // exp(-slant depth)) down to the absorption cut-off level and zero below
    // it, as a freshly constructed reference solver leaves it
    // (beam_transmission_ is only filled for lc < ncut before the cut-off
    // decision, src/FluxSolver.cpp:417-421).
    const double abst = s_abst;
    double bt_prev = (F0 > 0.0) ? 1.0 : 0.0;
    for (int lc = 0; lc < L; ++lc) {
      const int64_t g = lo + lc;
      double bt = 0.0;
      if (F0 > 0.0 && lc < ncut)
        bt = spherical ? exp(-a.lay_d[LD_CHTAU * TL + g]) : exp(-a.lev_d[VD_TAU_S * TV + vo + lc + 1] / mu0);
      a.lay_d[LD_BT_TOP * TL + g] = bt_prev;
      a.lay_d[LD_BT_BOT * TL + g] = bt;
      bt_prev = bt;
    }
    const bool lyrcut = (abst >= 10.0 && !thermal && L > 1 && !(a.flags & kFluxNoCutoff));
    const int ncut_f = lyrcut ? ncut : L;
    a.inst_i[i * kInstI + II_NCUT] = ncut_f;
    a.inst_i[i * kInstI + II_LYRCUT] = lyrcut ? 1 : 0;
    a.inst_i[i * kInstI + II_STATUS] = 0;
    a.inst_i[i * kInstI + II_SCALING] = (deltam ? 1 : 0) | (dmp ? 2 : 0);
    a.inst_d[i * kInstD + ID_BTRANS_NCUT] = (ncut_f == 0) ? bt_prev : a.lay_d[LD_BT_BOT * TL + lo + ncut_f - 1];
  }
  // Planck function per level (src/FluxSolver.cpp:104-122, :748-752).
  if (thermal && a.planck_ovr) {
    // Given Planck values (temperature-Jacobian solves).
    for (int l = lane; l <= L; l += 64) a.lev_d[VD_PLANCK * TV + vo + l] = a.planck_ovr[vo + l];
    if (lane == 0) {
      a.inst_d[i * kInstD + ID_PL_TOP] = a.planck_ovr[TV + 2 * i];
      const double pb = a.planck_ovr[TV + 2 * i + 1];
      a.inst_d[i * kInstD + ID_PL_BOT] = pb;
      double pbd = 0.0;
      if (a.flags & DISORT_F_DIFFUSION_BC) {
        const double dl = a.lay_d[LD_DTAU_S * TL + lo + L - 1];
        if (dl > 0.0) pbd = (pb - a.planck_ovr[vo + L - 1]) / dl;
      }
      a.inst_d[i * kInstD + ID_PL_BOT_DERIV] = pbd;
    }
  } else if (thermal) {
    const double wlo = sc[DISORT_S_WN_LO], whi = sc[DISORT_S_WN_HI];
    for (int l = lane; l <= L; l += 64) a.lev_d[VD_PLANCK * TV + vo + l] = planck_band(wlo, whi, a.temper[vo + l]);
    if (lane == 0) {
      a.inst_d[i * kInstD + ID_PL_TOP] = planck_band(wlo, whi, sc[DISORT_S_TEMP_TOP]) * sc[DISORT_S_EMIS_TOP];
      const double pb = planck_band(wlo, whi, sc[DISORT_S_TEMP_BOT]);
      a.inst_d[i * kInstD + ID_PL_BOT] = pb;
      double pbd = 0.0;
      if (a.flags & DISORT_F_DIFFUSION_BC) {
        const double bl = planck_band(wlo, whi, a.temper[vo + L - 1]);
        const double dl = a.lay_d[LD_DTAU_S * TL + lo + L - 1];
        if (dl > 0.0) pbd = (pb - bl) / dl;
      }
      a.inst_d[i * kInstD + ID_PL_BOT_DERIV] = pbd;
    }
  } else if (lane == 0) {
    a.inst_d[i * kInstD + ID_PL_TOP] = 0.0;
    a.inst_d[i * kInstD + ID_PL_BOT] = 0.0;
    a.inst_d[i * kInstD + ID_PL_BOT_DERIV] = 0.0;
  }
}


// Entry G(q, j) of the eigenvector matrix of task t (reference ordering),
// with a run-time row q.
template <int NN>
__device__ __forceinline__ double g_entry(const FluxArgs& a, int64_t t, int q, int j) {
  constexpr EigOff eo(2 * NN);
  const bool up = j >= NN;
  const int jp = up ? j - NN : NN - 1 - j;
  const bool lower = q >= NN;
  const int r = lower ? q - NN : NN - 1 - q;
  // lower & up: gp; upper & up: gm; lower & down: -gm; upper & down: -gp
  const int f = (lower == up) ? eo.gp() : eo.gm();
  const double v = a.eig[eo.at(t, f + r + jp * NN)];
  return up ? v : -v;
}

// ===========================================================================
// k_flux: fluxes and mean intensities at the layer boundaries.
//
// A group of LPL lanes per level (LPL = NN rounded up to a power of two),
// lane r owning the quadrature angle pair (+mu_r, -mu_r): it reads row r of
// the layer's gp/gm blocks (consecutive lanes, consecutive addresses), and
// the per-angle terms of the fluxes are summed over the group.
// ===========================================================================
template <int NN>
struct FluxLanes {
  static constexpr int LPL = NN <= 2 ? 2 : NN <= 4 ? 4 : NN <= 8 ? 8 : NN <= 16 ? 16 : 32;
  static constexpr int LPW = 64 / LPL;  // levels per wavefront
  static_assert(NN <= 32, "at most 32 angle pairs");
};

template <int NN>
__global__ __launch_bounds__(64) void k_flux(FluxArgs a) {
  constexpr int N = 2 * NN;
  constexpr EigOff eo(N);
  constexpr TabOff to(N);
  constexpr int LPL = FluxLanes<NN>::LPL, LPW = FluxLanes<NN>::LPW;
  __shared__ double sY[LPW * 2 * LPL];
  const int lv = threadIdx.x / LPL;
  const int r = threadIdx.x % LPL;
  const int64_t vt = (int64_t)blockIdx.x * LPW + lv;
  if (vt >= a.n_lvtasks) return;  // whole groups leave together
  const int bk = a.v_bk[vt];
  const int lu = a.v_lu[vt];
  const int i = a.b_inst[bk];
  const int64_t t0 = a.b_task0[bk];
  const int64_t lo = a.lay_off[i], vo = lo + i;
  const int64_t TL = a.total_layers, TV = a.total_levels, OV = a.total_levels;
  const double* sc = a.scal + (int64_t)i * DISORT_NSCAL;
  const double* tb = a.tab;
  const int ncut = a.inst_i[i * kInstI + II_NCUT];
  const bool lyrcut = a.inst_i[i * kInstI + II_LYRCUT] != 0;
  int lyu = (lu == 0) ? 0 : lu - 1;
  double o[DISORT_NOUT];
#pragma unroll
  for (int f = 0; f < DISORT_NOUT; ++f) o[f] = 0.0;
  if (!(lyrcut && lyu > ncut - 1)) {
    if (lyu >= ncut) lyu = ncut - 1;
    const int64_t tt = t0 + lyu;
    const double* rec = a.eig + eo.at(tt, 0);
    const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU];
    const double ut = a.lev_d[VD_TAU_S * TV + vo + lu];
    const double dts = a.lay_d[LD_DTAU_S * TL + lo + lyu];
    const bool act = r < NN;
    const int rr = act ? r : 0;
    // Mode amplitudes at this level, C_j e_j, one mode pair per lane (mode
    // NN-1-jp has eigenvalue -k_jp, mode NN+jp +k_jp): y[jp] downward,
    // y[LPL+jp] upward.
    double* y = sY + lv * 2 * LPL;
    {
      const double e = exp(-rec[eo.k() + rr] * dts);
      const double yd = a.cvec[tt * N + NN - 1 - rr] * ((lu == 0) ? e : 1.0);
      const double yu = a.cvec[tt * N + NN + rr] * ((lu == 0) ? 1.0 : e);
      wave_lds_order();
      if (act) {
        y[r] = yd;
        y[LPL + r] = yu;
      }
      wave_lds_order();
    }
    // Intensity at the quadrature angles: I = G (C o e) + Z.
    const int zoff = (lu == 0) ? eo.ztop() : eo.zfl();
    double iu = rec[zoff + NN + rr], id = rec[zoff + NN - 1 - rr];
    constexpr int kUnroll = NN <= 16 ? NN : 8;
#pragma unroll kUnroll
    for (int jp = 0; jp < NN; ++jp) {
      const double yd = y[jp], yu = y[LPL + jp];
      const double gp = rec[eo.gp() + rr + jp * NN];
      const double gm = rec[eo.gm() + rr + jp * NN];
      iu += gp * yu - gm * yd;
      id += gm * yu - gp * yd;
    }
    const double w = act ? tb[to.cwt() + rr] : 0.0, cm = tb[to.cmu() + rr];
    double mi = group_sum<LPL>(w * id + w * iu);
    double mid = group_sum<LPL>(w * id);
    double miu = group_sum<LPL>(w * iu);
    double fldn = group_sum<LPL>(w * id * cm);
    double flup = group_sum<LPL>(w * iu * cm);
    double fact = 0.0, dirint = 0.0, fldir = 0.0, fdir = 0.0;
    if (F0 > 0.0) {
      // Path factor mu0, or the layer's Chapman factor (:1184-1194).
      const double ch = a.lay_d[LD_CH * TL + lo + lyu];
      const double tauc = a.lev_d[VD_TAUC * TV + vo + lu];
      fact = exp(-ut / ch);
      fdir = fabs(mu0) * F0 * exp(-tauc / ch);
      dirint = F0 * fact;
      fldir = mu0 * F0 * fact;
    }
    flup *= 2.0 * kPi;
    fldn *= 2.0 * kPi;
    const double fdn = (fldn + fldir) - fdir;
    mi = (2.0 * kPi * mi + dirint) / (4.0 * kPi);
    const double midir = dirint / (4.0 * kPi);
    mid = (2.0 * kPi * mid) / (4.0 * kPi);
    miu = (2.0 * kPi * miu) / (4.0 * kPi);
    double pls = 0.0;
    if (a.flags & DISORT_F_THERMAL)
      pls = (lu == 0) ? a.lev_d[VD_PLANCK * TV + vo] : rec[eo.pls()];
    const double div = (1.0 - a.ssa[lo + lyu]) * 4.0 * kPi * (mi - pls);
    o[DISORT_O_FLUX_DIRECT_BEAM] = fdir;
    o[DISORT_O_FLUX_DOWN] = fdn;
    o[DISORT_O_FLUX_UP] = flup;
    o[DISORT_O_FLUX_TAU_DIVERGENCE] = div;
    o[DISORT_O_MEAN_INTENSITY] = mi;
    o[DISORT_O_MEAN_INTENSITY_DOWN] = mid;
    o[DISORT_O_MEAN_INTENSITY_UP] = miu;
    o[DISORT_O_MEAN_INTENSITY_DIRECT_BEAM] = midir;
  }
  if (r == 0) {
#pragma unroll
    for (int f = 0; f < DISORT_NOUT; ++f) a.out[f * OV + vo + lu] = o[f];
  }
}

// Launch of k_flux over the level tasks of a bucket.
template <int NN>
inline void launch_flux_levels(const FluxArgs& a, hipStream_t s) {
  constexpr int LPW = FluxLanes<NN>::LPW;
  hipLaunchKernelGGL(k_flux<NN>, dim3((unsigned)((a.n_lvtasks + LPW - 1) / LPW)), dim3(64), 0, s, a);
}

}  // namespace
}  // namespace dsa
