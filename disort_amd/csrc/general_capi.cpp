// Host side of disort_general_solve (include/disort_amd.h): validation with
// the messages of DisortConfig::validate (src/DisortConfig.cpp:72-249),
// constant tables (associated Legendre functions at the nodes and the user
// angles, the BRDF azimuth quadrature), de-duplication of RPV parameter
// sets, chunking of the columns to a scratch budget, and the launches of
// general_kernels.hip on the context stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/disort_amd.h"
#include "ctx.hpp"
#include "flux_args.hpp"
#include "flux_layout.hpp"
#include "general_layout.hpp"
#include "planck.hpp"

using namespace dsa_host;

namespace {

// Y_l^m(mu) for m = 0..N-1, l = 0..N-1 at nmu cosines, out[(m*N + l)*nmu + k],
// by the recurrence of src/Legendre.cpp:67-120 applied mode after mode (the
// diagonal Y_m^m is built from Y_{m-1}^{m-1} of the previous mode, as the
// reference's solve loop does).
void legendre_modes(const double* mu, int nmu, int N, std::vector<double>& out) {
  std::vector<double> y((size_t)(N + 1) * nmu, 0.0);  // rows l = 0..N
  out.assign((size_t)N * N * nmu, 0.0);
  for (int m = 0; m < N; ++m) {
    if (m == 0) {
      for (int k = 0; k < nmu; ++k) {
        y[k] = 1.0;
        y[nmu + k] = mu[k];
      }
      for (int l = 2; l <= N - 1; ++l)
        for (int k = 0; k < nmu; ++k)
          y[(size_t)l * nmu + k] = ((2.0 * l - 1.0) * mu[k] * y[(size_t)(l - 1) * nmu + k] -
                                    (l - 1.0) * y[(size_t)(l - 2) * nmu + k]) / l;
    } else {
      for (int k = 0; k < nmu; ++k) {
        y[(size_t)m * nmu + k] = -std::sqrt((1.0 - 1.0 / (2.0 * m)) * (1.0 - mu[k] * mu[k])) * y[(size_t)(m - 1) * nmu + k];
        y[(size_t)(m + 1) * nmu + k] = std::sqrt(2.0 * m + 1.0) * mu[k] * y[(size_t)m * nmu + k];
      }
      for (int l = m + 2; l <= N - 1; ++l) {
        const double a = std::sqrt((double)((l - m) * (l + m)));
        const double b = std::sqrt((double)((l - m - 1) * (l + m - 1)));
        for (int k = 0; k < nmu; ++k)
          y[(size_t)l * nmu + k] = ((2.0 * l - 1.0) * mu[k] * y[(size_t)(l - 1) * nmu + k] - b * y[(size_t)(l - 2) * nmu + k]) / a;
      }
    }
    for (int l = m; l < N; ++l)
      for (int k = 0; k < nmu; ++k) out[((size_t)m * N + l) * nmu + k] = y[(size_t)l * nmu + k];
  }
}

// [ylm | ylq | ylw] for N streams: Y_l^m at the positive nodes, scaled by
// sqrt(w/mu), and times the quadrature weight w, cached per stream count.
int mode_tables(disort_ctx* ctx, int N, double** out) {
  auto it = ctx->mode_tables.find(N);
  if (it != ctx->mode_tables.end()) {
    *out = it->second;
    return DISORT_OK;
  }
  const int NN = N / 2;
  std::vector<double> cmu, cwt;
  if (!gauss_nodes(NN, cmu, cwt)) return fail(ctx, DISORT_E_NUMERIC, "gaussianQuadrature: Failed to converge");
  std::vector<double> y;
  legendre_modes(cmu.data(), NN, N, y);
  std::vector<double> t(3 * y.size());
  for (size_t e = 0; e < y.size(); ++e) {
    const int r = (int)(e % NN);
    t[e] = y[e];
    t[y.size() + e] = y[e] * std::sqrt(cwt[r] / cmu[r]);
    t[2 * y.size() + e] = cwt[r] * y[e];
  }
  double* d = nullptr;
  HIPCHK(ctx, hipMalloc(&d, t.size() * sizeof(double)));
  HIPCHK(ctx, hipMemcpy(d, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
  ctx->mode_tables[N] = d;
  *out = d;
  return DISORT_OK;
}

// One column against src/DisortConfig.cpp:72-249 (messages verbatim) and
// the limits of the GPU path.  Returns "" when valid; *code receives the
// error class.
std::string validate_column(const disort_general_batch* b, int64_t i, int* code) {
  using std::to_string;
  *code = DISORT_E_INVALID;
  const int L = b->nlyr, P = b->pstride, N = b->nstr;
  const uint32_t f = (uint32_t)b->flags;
  const bool thermal = f & DISORT_F_THERMAL, diffusion = f & DISORT_F_DIFFUSION_BC;
  const bool brdf = f & dsa::kBrdfFlags, user_mu = f & DISORT_G_USER_MU, only = f & DISORT_G_ONLY_FLUXES;
  const double* s = b->scal + i * DISORT_NSCAL;
  if (L <= 0) return "num_layers must be positive";
  if (N < 4) return "num_streams must be >= 4";
  if (N % 2 != 0) return "num_streams must be even";
  if (P < N + 1) return "phase_function_moments[" + to_string(0) + "] must have at least nmomNstr()+1 elements";
  for (int lc = 0; lc < L; ++lc) {
    if (b->dtau[i * L + lc] < 0.0) return "delta_tau[" + to_string(lc) + "] is negative";
    const double w = b->ssa[i * L + lc];
    if (w < 0.0 || w > 1.0) return "single_scat_albedo[" + to_string(lc) + "] must be in [0, 1]";
  }
  if (thermal) {
    if (!b->temper) return "temperature array is too small (need num_layers+1 when use_thermal_emission=true)";
    for (int l = 0; l <= L; ++l)
      if (b->temper[i * (L + 1) + l] < 0.0) return "temperature[" + to_string(l) + "] is negative";
    if (s[DISORT_S_WN_LO] < 0.0 || s[DISORT_S_WN_HI] < s[DISORT_S_WN_LO])
      return "Invalid wavenumber range for Planck function";
  }
  if (diffusion) {
    if (!thermal) return "use_diffusion_lower_bc requires use_thermal_emission = true";
    if (brdf) return "use_diffusion_lower_bc requires use_lambertian_surface = true (BRDF is incompatible)";
  }
  const double F0 = s[DISORT_S_BEAM_FLUX], mu0 = s[DISORT_S_BEAM_MU];
  if (F0 < 0.0) return "direct_beam_flux must be non-negative";
  if ((mu0 <= 0.0 || mu0 > 1.0) && F0 > 0.0) return "direct_beam_mu must be in (0, 1] when direct_beam_flux > 0";
  if (!brdf && (s[DISORT_S_ALBEDO] < 0.0 || s[DISORT_S_ALBEDO] > 1.0))
    return "surface_albedo must be in [0, 1] when use_lambertian_surface=true";
  if (user_mu) {
    if (b->numu <= 0) return "num_user_mu must be positive when use_user_mu=true";
    for (int iu = 0; iu < b->numu; ++iu) {
      if (std::abs(b->mu_user[iu]) > 1.0) return "mu_user[" + to_string(iu) + "] out of range [-1, 1]";
      if (iu > 0 && b->mu_user[iu] <= b->mu_user[iu - 1]) return "mu_user must be in increasing order";
    }
  }
  if ((f & DISORT_F_SPHERICAL)) {
    if (s[DISORT_S_RADIUS] <= 0.0) return "bottom_radius must be positive when use_spherical_beam=true";
    if (!b->alt) return "level_altitudes array is too small (need num_layers+1 when use_spherical_beam=true)";
  }
  if (!only) {
    if (b->nphi <= 0) return "num_phi must be positive when comp_only_fluxes=false";
    for (int j = 0; j < b->nphi; ++j)
      if (b->phi_user[j] < 0.0 || b->phi_user[j] > 360.0) return "phi_user[" + to_string(j) + "] must be in [0, 360]";
  }
  if ((f & DISORT_G_CORR_BURAS) && !(f & DISORT_G_CORR_NAKAJIMA) && b->nphase > 0) {
    if (b->nphase < 2 || b->nphase > dsa::kMaxPhaseAngles)
      return "num_phase_func_angles must be in [2, " + to_string(dsa::kMaxPhaseAngles) + "]";
    if (!b->mu_phase || !b->phase) return "mu_phase_function and phase_function must be set when num_phase_func_angles > 0";
    for (int it = 0; it < b->nphase; ++it) {
      if (std::abs(b->mu_phase[it]) > 1.0) return "mu_phase_function[" + to_string(it) + "] out of range [-1, 1]";
      if (it > 0 && b->mu_phase[it] <= b->mu_phase[it - 1]) return "mu_phase_function must be in increasing order";
    }
  }
  if ((f & DISORT_F_DELTA_M_PLUS) && P - 1 < N + 1)
    return "use_delta_m_plus=true requires num_phase_func_moments >= num_streams + 1 (need moments M and M+1)";
  if (b->accuracy < 0.0) return "accuracy_fourier_series must be non-negative (0 = disabled)";
  if (thermal) {
    if (s[DISORT_S_WN_LO] < 0.0 || s[DISORT_S_WN_HI] <= s[DISORT_S_WN_LO])
      return "use_thermal_emission=true requires 0 <= wavenumber_low < wavenumber_high";
    if (s[DISORT_S_TEMP_TOP] < 0.0 || s[DISORT_S_TEMP_BOT] < 0.0)
      return "use_thermal_emission=true requires non-negative temperature_top and temperature_bottom";
    if (s[DISORT_S_EMIS_TOP] < 0.0 || s[DISORT_S_EMIS_TOP] > 1.0)
      return "use_thermal_emission=true requires 0 <= emissivity_top <= 1";
  }
  *code = DISORT_E_UNSUPPORTED;
  if (!gpu_supported_nstr(N)) return "Unsupported num_streams=" + to_string(N) + " (GPU general path: 4..16 even, 32, 64)";
  return "";
}

struct Dev {
  std::vector<void*> owned;
  ~Dev() {
    for (void* p : owned) (void)hipFree(p);
  }
};

}  // namespace

namespace {
// The general solve; with ovr (host, [n (nlyr+1)] level Planck values then
// [n][2] top emission and B(T_bottom)) the Planck values are given instead
// of computed from the temperatures (the temperature-Jacobian solves).
int general_solve(disort_ctx* ctx, const disort_general_batch* b, int memspace, const double* ovr) {
  using namespace dsa;
  if (!ctx || !b) return fail(ctx, DISORT_E_INVALID, "null argument");
  if (b->n <= 0) return DISORT_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int N = b->nstr, NN = N / 2, L = b->nlyr, P = b->pstride;
  const uint32_t f = (uint32_t)b->flags;
  const bool host = memspace == DISORT_MEM_HOST;
  const bool user_tau = f & DISORT_G_USER_TAU, only = f & DISORT_G_ONLY_FLUXES;
  const bool do_user = (f & DISORT_G_USER_MU) && !only;
  const bool brdf = f & dsa::kBrdfFlags;
  const int32_t bflags = f & dsa::kBrdfFlags;
  if (bflags & (bflags - 1)) return fail(ctx, DISORT_E_INVALID, "more than one BRDF type selected");
  const int brdf_kind = (bflags == DISORT_G_BRDF_HAPKE)     ? dsa::kBrdfHapke
                        : (bflags == DISORT_G_BRDF_COXMUNK) ? dsa::kBrdfCoxMunk
                        : (bflags == DISORT_G_BRDF_ROSSLI)  ? dsa::kBrdfRossLi
                                                            : dsa::kBrdfRpv;
  const int64_t n = b->n;
  const int ntau = user_tau ? b->ntau : L + 1;
  const int numu = do_user ? b->numu : N;
  const int nphi = only ? 0 : b->nphi;
  if (ntau != b->ntau) return fail(ctx, DISORT_E_INVALID, "ntau must be nlyr+1 without DISORT_G_USER_TAU");
  if (numu != b->numu) return fail(ctx, DISORT_E_INVALID, "numu must be nstr without DISORT_G_USER_MU");
  if (user_tau && (ntau <= 0 || !b->tau_user)) return fail(ctx, DISORT_E_INVALID, "use_user_tau needs tau_user");
  if (do_user && !b->mu_user) return fail(ctx, DISORT_E_INVALID, "use_user_mu needs mu_user");
  if (!only && !b->phi_user) return fail(ctx, DISORT_E_INVALID, "phi_user array is too small");
  if (brdf && !b->rpv) return fail(ctx, DISORT_E_INVALID, "bidirReflectivity: RPV BRDF selected but rpv spec is null");
  if (!gpu_supported_nstr(N))
    return fail(ctx, DISORT_E_UNSUPPORTED, "Unsupported num_streams=" + std::to_string(N) + " (GPU general path: 4..16 even, 32, 64)");
  if (host)
    for (int64_t i = 0; i < n; ++i) {
      int code;
      const std::string m = validate_column(b, i, &code);
      if (!m.empty()) return fail(ctx, code, "column " + std::to_string(i) + ": " + m);
    }

  // Small host copies: user angles, azimuths, RPV parameters.
  std::vector<double> mu_h, phi_h, rpv_h;
  auto fetch = [&](const double* src, size_t cnt, std::vector<double>& dst) -> int {
    dst.resize(cnt);
    if (!cnt) return DISORT_OK;
    if (host) std::memcpy(dst.data(), src, cnt * 8);
    else HIPCHK(ctx, hipMemcpy(dst.data(), src, cnt * 8, hipMemcpyDeviceToHost));
    return DISORT_OK;
  };
  int rc;
  if (do_user && (rc = fetch(b->mu_user, numu, mu_h))) return rc;
  if (nphi && (rc = fetch(b->phi_user, nphi, phi_h))) return rc;
  if (brdf && (rc = fetch(b->rpv, (size_t)n * DISORT_NRPV, rpv_h))) return rc;

  // RPV parameter sets (columns usually share one surface).
  std::vector<int32_t> set_of(n, 0);
  std::vector<double> sets;
  if (brdf) {
    std::map<std::vector<double>, int> idx;
    for (int64_t i = 0; i < n; ++i) {
      std::vector<double> key(rpv_h.begin() + i * DISORT_NRPV, rpv_h.begin() + (i + 1) * DISORT_NRPV);
      auto it = idx.find(key);
      if (it == idx.end()) {
        it = idx.emplace(key, (int)idx.size()).first;
        sets.insert(sets.end(), key.begin(), key.end());
      }
      set_of[i] = it->second;
    }
  }
  const int nsets = brdf ? (int)(sets.size() / DISORT_NRPV) : 0;
  const int nm = (do_user && !only) ? N : 1;

  // Device inputs (host memspace: uploaded into context buffers).
  GenArgs a{};
  a.flags = b->flags;
  a.nstr = N;
  a.nlyr = L;
  a.pstride = P;
  a.ntau = ntau;
  a.numu = numu;
  a.nphi = nphi;
  a.nm = nm;
  a.n = n;
  a.accuracy = b->accuracy;
  struct Up {
    const char* name;
    const void* src;
    size_t bytes;
    const double** dst;
  };
  const bool thermal = f & DISORT_F_THERMAL, spherical = f & DISORT_F_SPHERICAL;
  std::vector<Up> ups = {{"g_dtau", b->dtau, (size_t)n * L * 8, &a.dtau},
                         {"g_ssa", b->ssa, (size_t)n * L * 8, &a.ssa},
                         {"g_pmom", b->pmom, (size_t)n * L * P * 8, &a.pmom},
                         {"g_scal", b->scal, (size_t)n * DISORT_NSCAL * 8, &a.scal}};
  if (thermal) ups.push_back({"g_temp", b->temper, (size_t)n * (L + 1) * 8, &a.temper});
  if (spherical) ups.push_back({"g_alt", b->alt, (size_t)n * (L + 1) * 8, &a.alt});
  if (user_tau) ups.push_back({"g_utau", b->tau_user, (size_t)n * ntau * 8, &a.tau_user});
  if (do_user) ups.push_back({"g_umu", b->mu_user, (size_t)numu * 8, &a.mu_user});
  if (nphi) ups.push_back({"g_phi", b->phi_user, (size_t)nphi * 8, &a.phi_user});
  // Buras-Emde correction: a tabulated phase function with
  // intensity_corr_buras and without intensity_corr_nakajima
  // (src/DisortSolver.cpp:369-377).
  const bool buras = do_user && (f & DISORT_G_CORR_BURAS) && !(f & DISORT_G_CORR_NAKAJIMA) && b->nphase > 0;
  const int np = buras ? b->nphase : 0;
  if (buras) {
    if (np < 2 || np > dsa::kMaxPhaseAngles || !b->mu_phase || !b->phase)
      return fail(ctx, DISORT_E_INVALID, "tabulated phase function: need 2..2048 angles with mu_phase and phase");
    ups.push_back({"g_muph", b->mu_phase, (size_t)np * 8, &a.mu_phase});
    ups.push_back({"g_phase", b->phase, (size_t)n * L * np * 8, &a.phase});
  }
  a.nphase = np;
  for (auto& u : ups) {
    if (host) {
      void* p;
      if ((rc = get_buf(ctx, u.name, u.bytes, &p))) return rc;
      HIPCHK(ctx, hipMemcpyAsync(p, u.src, u.bytes, hipMemcpyHostToDevice, ctx->stream));
      *u.dst = (const double*)p;
    } else {
      *u.dst = (const double*)u.src;
    }
  }
  const size_t flux_bytes = (size_t)DISORT_NOUT * n * ntau * 8;
  const size_t inten_bytes = (size_t)n * ntau * numu * nphi * 8, avg_bytes = (size_t)n * ntau * numu * 8;
  void* p;
  if (host) {
    if ((rc = get_buf(ctx, "g_flux", flux_bytes, &p))) return rc;
    a.flux = (double*)p;
    if ((rc = get_buf(ctx, "g_inten", inten_bytes, &p))) return rc;
    a.inten = (double*)p;
    if ((rc = get_buf(ctx, "g_avg", avg_bytes, &p))) return rc;
    a.inten_avg = (double*)p;
    if ((rc = get_buf(ctx, "g_status", (size_t)n * 4, &p))) return rc;
    a.status = (int32_t*)p;
  } else {
    a.flux = b->flux;
    a.inten = b->inten;
    a.inten_avg = b->inten_avg;
    if (!b->status) {
      if ((rc = get_buf(ctx, "g_status", (size_t)n * 4, &p))) return rc;
      a.status = (int32_t*)p;
    } else {
      a.status = b->status;
    }
  }
  if (nphi && !do_user && a.inten) HIPCHK(ctx, hipMemsetAsync(a.inten, 0, inten_bytes, ctx->stream));
  if (!do_user && a.inten_avg) HIPCHK(ctx, hipMemsetAsync(a.inten_avg, 0, avg_bytes, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(a.status, 0, (size_t)n * 4, ctx->stream));

  // Batch-wide per-column records (k_column through a flux-batch view).
  const int64_t TL = n * L, TV = n * (L + 1);
  std::vector<int32_t> vn(n, N), vl(n, L), vp(n, P), bi(n);
  std::vector<int64_t> loff(n), moff(n);
  for (int64_t i = 0; i < n; ++i) {
    bi[i] = (int32_t)i;
    loff[i] = i * L;
    moff[i] = i * L * P;
  }
  FluxArgs fa{};
  struct UpI {
    const char* name;
    const void* src;
    size_t bytes;
    void** dst;
  };
  void *d_vn, *d_vl, *d_vp, *d_bi, *d_lo, *d_mo, *d_set = nullptr, *d_sets = nullptr;
  std::vector<UpI> upi = {{"g_vn", vn.data(), (size_t)n * 4, &d_vn}, {"g_vl", vl.data(), (size_t)n * 4, &d_vl},
                          {"g_vp", vp.data(), (size_t)n * 4, &d_vp}, {"g_bi", bi.data(), (size_t)n * 4, &d_bi},
                          {"g_lo", loff.data(), (size_t)n * 8, &d_lo}, {"g_mo", moff.data(), (size_t)n * 8, &d_mo}};
  if (brdf) {
    upi.push_back({"g_set", set_of.data(), (size_t)n * 4, &d_set});
    upi.push_back({"g_sets", sets.data(), sets.size() * 8, &d_sets});
  }
  for (auto& u : upi) {
    if ((rc = get_buf(ctx, u.name, u.bytes, u.dst))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(*u.dst, u.src, u.bytes, hipMemcpyHostToDevice, ctx->stream));
  }
  void *d_layd, *d_levd, *d_instd, *d_insti, *d_ulyr, *d_utau;
  if ((rc = get_buf(ctx, "g_layd", (size_t)kLayD * TL * 8, &d_layd))) return rc;
  if ((rc = get_buf(ctx, "g_levd", (size_t)kLevD * TV * 8, &d_levd))) return rc;
  if ((rc = get_buf(ctx, "g_instd", (size_t)n * kInstD * 8, &d_instd))) return rc;
  if ((rc = get_buf(ctx, "g_insti", (size_t)n * kInstI * 4, &d_insti))) return rc;
  if ((rc = get_buf(ctx, "g_ulyr", (size_t)n * ntau * 4, &d_ulyr))) return rc;
  if ((rc = get_buf(ctx, "g_utaus", (size_t)n * ntau * 8, &d_utau))) return rc;
  fa.flags = b->flags & (DISORT_F_THERMAL | DISORT_F_SPHERICAL | DISORT_F_DELTA_M_PLUS | DISORT_F_DIFFUSION_BC);
  if (b->flags & DISORT_G_NO_CUTOFF) fa.flags |= kFluxNoCutoff;
  if (ovr) {
    void* d_ovr;
    const size_t ob = (size_t)(TV + 2 * n) * 8;
    if ((rc = get_buf(ctx, "g_ovr", ob, &d_ovr))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(d_ovr, ovr, ob, hipMemcpyHostToDevice, ctx->stream));
    fa.planck_ovr = (const double*)d_ovr;
  }
  fa.nstr = (const int32_t*)d_vn;
  fa.nlyr = (const int32_t*)d_vl;
  fa.pstride = (const int32_t*)d_vp;
  fa.lay_off = (const int64_t*)d_lo;
  fa.mom_off = (const int64_t*)d_mo;
  fa.dtau = a.dtau;
  fa.ssa = a.ssa;
  fa.pmom = a.pmom;
  fa.temper = a.temper;
  fa.alt = a.alt;
  fa.scal = a.scal;
  fa.total_levels = TV;
  fa.total_layers = TL;
  fa.inst_d = (double*)d_instd;
  fa.inst_i = (int32_t*)d_insti;
  fa.lay_d = (double*)d_layd;
  fa.lev_d = (double*)d_levd;
  fa.nb = (int32_t)n;
  fa.b_inst = (const int32_t*)d_bi;
  a.lay_d = fa.lay_d;
  a.lev_d = fa.lev_d;
  a.inst_d = fa.inst_d;
  a.inst_i = fa.inst_i;
  a.u_lyr = (int32_t*)d_ulyr;
  a.u_tau = (double*)d_utau;
  a.rpv_set = (const int32_t*)d_set;
  a.set_rpv = (const double*)d_sets;
  a.nsets = nsets;
  a.brdf_kind = brdf_kind;

  // Constant tables.
  double *tab, *mt;
  if ((rc = table_for(ctx, N, &tab))) return rc;
  if ((rc = mode_tables(ctx, N, &mt))) return rc;
  a.tab = tab;
  a.ylm = mt;
  a.ylq = mt + (size_t)N * N * NN;
  a.ylw = mt + 2 * (size_t)N * N * NN;
  std::vector<double> ylu_h;
  if (do_user) legendre_modes(mu_h.data(), numu, N, ylu_h);
  std::vector<double> gaz_h(2 * 200, 0.0);
  {
    std::vector<double> gm, gw;
    if (!gauss_nodes(100, gm, gw)) return fail(ctx, DISORT_E_NUMERIC, "gaussianQuadrature: Failed to converge");
    std::copy(gm.begin(), gm.end(), gaz_h.begin());
    std::copy(gw.begin(), gw.end(), gaz_h.begin() + 200);
  }
  void *d_ylu = nullptr, *d_gaz;
  if (do_user) {
    if ((rc = get_buf(ctx, "g_ylu", ylu_h.size() * 8, &d_ylu))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(d_ylu, ylu_h.data(), ylu_h.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  }
  if ((rc = get_buf(ctx, "g_gaz", gaz_h.size() * 8, &d_gaz))) return rc;
  HIPCHK(ctx, hipMemcpyAsync(d_gaz, gaz_h.data(), gaz_h.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  a.ylu = (const double*)d_ylu;
  a.gaz = (const double*)d_gaz;
  if (brdf) {
    if ((rc = get_buf(ctx, "g_srfd", (size_t)nsets * nm * surf_d_size(NN, numu) * 8, &p))) return rc;
    a.srf_d = (double*)p;
  }

  // Chunk size from a scratch budget.
  const GE ge(N);
  const size_t per_cm = (size_t)L * (ge.size() + (size_t)numu * N + (size_t)kUSrc * numu + panel_stride(N) + N +
                                     (size_t)4 * numu) * 8 +
                        (size_t)ntau * numu * 8 + (size_t)surf_b_size(NN, numu) * 8;
  size_t free_b = 0, total_b = 0;
  HIPCHK(ctx, hipMemGetInfo(&free_b, &total_b));
  const size_t budget = std::max<size_t>(std::min<size_t>(free_b / 2, (size_t)24 << 30), (size_t)256 << 20);
  const size_t be_per_col = buras ? (size_t)ntau * ((size_t)np * 8 + (size_t)dsa::kBeNF * 12 + 8) : 0;
  int64_t nc = (int64_t)std::max<size_t>(1, budget / (per_cm * nm + be_per_col));
  nc = std::min<int64_t>(nc, n);
  const int64_t nt_max = nc * nm * L;
  void *d_eig, *d_gu, *d_us, *d_pan, *d_cv, *d_uum, *d_uc, *d_srfb, *d_modes;
  if ((rc = get_buf(ctx, "g_eig", (size_t)nt_max * ge.size() * 8, &d_eig))) return rc;
  if ((rc = get_buf(ctx, "g_gu", (size_t)nt_max * numu * N * 8, &d_gu))) return rc;
  if ((rc = get_buf(ctx, "g_us", (size_t)nt_max * kUSrc * numu * 8, &d_us))) return rc;
  if ((rc = get_buf(ctx, "g_pan", (size_t)nt_max * panel_stride(N) * 8, &d_pan))) return rc;
  if ((rc = get_buf(ctx, "g_cv", (size_t)nt_max * N * 8, &d_cv))) return rc;
  if ((rc = get_buf(ctx, "g_uum", (size_t)nc * nm * ntau * numu * 8, &d_uum))) return rc;
  if ((rc = get_buf(ctx, "g_uc", (size_t)nt_max * 4 * numu * 8, &d_uc))) return rc;
  if ((rc = get_buf(ctx, "g_srfb", (size_t)nc * nm * surf_b_size(NN, numu) * 8, &d_srfb))) return rc;
  if ((rc = get_buf(ctx, "g_modes", (size_t)nc * 4, &d_modes))) return rc;
  a.eig = (double*)d_eig;
  a.gu = (double*)d_gu;
  a.usrc = (double*)d_us;
  a.panel = (double*)d_pan;
  a.cvec = (double*)d_cv;
  a.uum = (double*)d_uum;
  a.srf_b = (double*)d_srfb;
  if (buras) {
    void *d_ph2, *d_grid, *d_neg, *d_norm;
    if ((rc = get_buf(ctx, "g_be_ph2", (size_t)nc * ntau * np * 8, &d_ph2))) return rc;
    if ((rc = get_buf(ctx, "g_be_grid", (size_t)nc * ntau * dsa::kBeNF * 8, &d_grid))) return rc;
    if ((rc = get_buf(ctx, "g_be_neg", (size_t)nc * ntau * dsa::kBeNF * 4, &d_neg))) return rc;
    if ((rc = get_buf(ctx, "g_be_norm", (size_t)nc * ntau * 8, &d_norm))) return rc;
    a.be_ph2 = (double*)d_ph2;
    a.be_grid = (double*)d_grid;
    a.be_neg = (int32_t*)d_neg;
    a.be_norm = (double*)d_norm;
  }

  const bool timed = ctx->timing;
  for (double& v : ctx->gstage_ms) v = 0.0;
  ctx->n_stages = 8;
  if (timed) HIPCHK(ctx, hipEventRecord(ctx->gev[0], ctx->stream));
  if (!general_setup(N, fa, a, ctx->stream, brdf)) return fail(ctx, DISORT_E_UNSUPPORTED, "no kernels for this stream count");
  HIPCHK(ctx, hipGetLastError());
  if (timed) {
    HIPCHK(ctx, hipEventRecord(ctx->gev[1], ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(ctx->gev[1]));
    float ms = 0;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->gev[0], ctx->gev[1]));
    ctx->gstage_ms[0] += ms;
  }
  GenLaunch g{};
  g.ucoef = (double*)d_uc;
  g.modes_used = (int32_t*)d_modes;
  g.brdf = brdf;
  g.do_user = do_user;
  // The intensity correction runs whenever intensity_corr_buras is set (the
  // kernels apply the rest of the gate per column): Buras-Emde with a
  // tabulated phase function and without intensity_corr_nakajima,
  // Nakajima-Tanaka otherwise (src/DisortSolver.cpp:364-377).
  g.nt = do_user && (f & DISORT_G_CORR_BURAS);
  g.buras = buras;
  g.brdf_corr = do_user && brdf && (f & DISORT_G_CORR_BURAS);
  g.ev = timed ? ctx->gev + 1 : nullptr;
  for (int64_t c0 = 0; c0 < n; c0 += nc) {
    a.c0 = c0;
    a.nc = std::min<int64_t>(nc, n - c0);
    a.n_tasks = a.nc * nm * L;
    if (!general_chunk(N, a, g, ctx->stream)) return fail(ctx, DISORT_E_UNSUPPORTED, "no kernels for this stream count");
    HIPCHK(ctx, hipGetLastError());
    if (timed) {
      HIPCHK(ctx, hipEventSynchronize(g.ev[do_user ? 8 : 4]));
      const int last = do_user ? 8 : 4;
      for (int k = 0; k < last; ++k) {
        float ms = 0;
        HIPCHK(ctx, hipEventElapsedTime(&ms, g.ev[k], g.ev[k + 1]));
        ctx->gstage_ms[1 + std::min(k, 6)] += ms;
      }
    }
  }
  if (!host) return DISORT_OK;
  std::vector<int32_t> st(n);
  HIPCHK(ctx, hipMemcpyAsync(b->flux, a.flux, flux_bytes, hipMemcpyDeviceToHost, ctx->stream));
  if (b->inten && nphi) HIPCHK(ctx, hipMemcpyAsync(b->inten, a.inten, inten_bytes, hipMemcpyDeviceToHost, ctx->stream));
  if (b->inten_avg) HIPCHK(ctx, hipMemcpyAsync(b->inten_avg, a.inten_avg, avg_bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(st.data(), a.status, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  int64_t bad = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (b->status) b->status[i] = st[i];
    bad += st[i] != 0;
  }
  if (bad) return fail(ctx, DISORT_E_NUMERIC, std::to_string(bad) + " column(s) failed (see status)");
  return DISORT_OK;
}
}  // namespace

extern "C" int disort_general_solve(disort_ctx* ctx, const disort_general_batch* b, int memspace) {
  return general_solve(ctx, b, memspace, nullptr);
}

// ---------------------------------------------------------------------------
// Temperature Jacobians of the general solver's fluxes
// (DisortSolver::computeThermalJacobian, src/DisortSolver.cpp:2130-2377).
// As for the flux path: the outputs are affine in the level Planck values
// and in B(T_bottom), so each degree of freedom is one column of an
// expanded flux-only batch with a unit Planck source there and no other
// source (no beam, no isotropic or top emission), scaled by dB/dT.  The
// surface emissivity (Lambertian or BRDF) and the diffusion lower boundary
// enter exactly as in the forward solve.
// ---------------------------------------------------------------------------
extern "C" int disort_general_jacobian(disort_ctx* ctx, const disort_general_batch* b, double* jac) {
  using namespace dsa;
  if (!ctx || !b || !jac) return fail(ctx, DISORT_E_INVALID, "null argument");
  if (!(b->flags & DISORT_F_THERMAL))
    return fail(ctx, DISORT_E_INVALID, "temperature Jacobians need use_thermal_emission = true");
  if (b->n <= 0) return DISORT_OK;
  const int L = b->nlyr, P = b->pstride, nd = L + 2;
  const bool user_tau = b->flags & DISORT_G_USER_TAU;
  const int ntau = user_tau ? b->ntau : L + 1;
  if (ntau != b->ntau) return fail(ctx, DISORT_E_INVALID, "ntau must be nlyr+1 without DISORT_G_USER_TAU");
  if (user_tau && (ntau <= 0 || !b->tau_user)) return fail(ctx, DISORT_E_INVALID, "use_user_tau needs tau_user");
  const bool brdf = b->flags & kBrdfFlags;
  if (brdf && !b->rpv) return fail(ctx, DISORT_E_INVALID, "bidirReflectivity: RPV BRDF selected but rpv spec is null");
  for (int64_t i = 0; i < b->n; ++i) {
    int code;
    const std::string m = validate_column(b, i, &code);
    if (!m.empty()) return fail(ctx, code, "column " + std::to_string(i) + ": " + m);
  }
  const int32_t flags = (b->flags & (DISORT_F_THERMAL | DISORT_F_DELTA_M_PLUS | DISORT_F_DIFFUSION_BC |
                                     DISORT_G_USER_TAU | DISORT_G_NO_CUTOFF | kBrdfFlags)) |
                        DISORT_G_ONLY_FLUXES;
  static const int field[6] = {DISORT_O_FLUX_UP, DISORT_O_FLUX_DOWN, DISORT_O_MEAN_INTENSITY,
                               DISORT_O_MEAN_INTENSITY_DOWN, DISORT_O_MEAN_INTENSITY_UP,
                               DISORT_O_FLUX_TAU_DIVERGENCE};
  const int64_t per = (int64_t)6 * ntau * nd;
  const int64_t cols = std::max<int64_t>(1, 16384 / nd);  // columns per expanded batch
  for (int64_t c0 = 0; c0 < b->n; c0 += cols) {
    const int64_t nc = std::min<int64_t>(cols, b->n - c0), ne = nc * nd;
    std::vector<double> dtau((size_t)ne * L), ssa((size_t)ne * L), pmom((size_t)ne * L * P),
        temper((size_t)ne * (L + 1)), scal((size_t)ne * DISORT_NSCAL), ovr((size_t)ne * (L + 1) + 2 * ne, 0.0),
        rpv(brdf ? (size_t)ne * DISORT_NRPV : 0), tau(user_tau ? (size_t)ne * ntau : 0);
    for (int64_t e = 0; e < ne; ++e) {
      const int64_t i = c0 + e / nd;
      const int k = (int)(e % nd);
      std::memcpy(&dtau[e * L], b->dtau + i * L, (size_t)L * 8);
      std::memcpy(&ssa[e * L], b->ssa + i * L, (size_t)L * 8);
      std::memcpy(&pmom[e * L * P], b->pmom + i * L * P, (size_t)L * P * 8);
      std::memcpy(&temper[e * (L + 1)], b->temper + i * (L + 1), (size_t)(L + 1) * 8);
      double* sc = &scal[e * DISORT_NSCAL];
      std::memcpy(sc, b->scal + i * DISORT_NSCAL, DISORT_NSCAL * 8);
      sc[DISORT_S_BEAM_FLUX] = 0.0;
      sc[DISORT_S_ISO_TOP] = 0.0;
      sc[DISORT_S_ISO_BOT] = 0.0;
      if (brdf) std::memcpy(&rpv[e * DISORT_NRPV], b->rpv + i * DISORT_NRPV, DISORT_NRPV * 8);
      if (user_tau) std::memcpy(&tau[e * ntau], b->tau_user + i * ntau, (size_t)ntau * 8);
      if (k <= L) ovr[e * (L + 1) + k] = 1.0;
      else ovr[ne * (L + 1) + 2 * e + 1] = 1.0;
    }
    std::vector<double> flux((size_t)DISORT_NOUT * ne * ntau);
    std::vector<int32_t> st(ne);
    disort_general_batch x{};
    x.n = ne;
    x.flags = flags;
    x.nstr = b->nstr;
    x.nlyr = L;
    x.pstride = P;
    x.ntau = ntau;
    x.numu = b->nstr;
    x.nphi = 0;
    x.accuracy = b->accuracy;
    x.dtau = dtau.data();
    x.ssa = ssa.data();
    x.pmom = pmom.data();
    x.temper = temper.data();
    x.scal = scal.data();
    x.rpv = brdf ? rpv.data() : nullptr;
    x.tau_user = user_tau ? tau.data() : nullptr;
    x.flux = flux.data();
    x.status = st.data();
    int rc = general_solve(ctx, &x, DISORT_MEM_HOST, ovr.data());
    if (rc) {
      for (int64_t e = 0; e < ne; ++e)
        if (st[e]) return fail(ctx, DISORT_E_NUMERIC, "column " + std::to_string(c0 + e / nd) + ": Jacobian solve failed");
      return rc;
    }
    for (int64_t e = 0; e < ne; ++e) {
      const int64_t i = c0 + e / nd;
      const int k = (int)(e % nd);
      const double* s2 = b->scal + i * DISORT_NSCAL;
      const double t = (k <= L) ? b->temper[i * (L + 1) + k] : s2[DISORT_S_TEMP_BOT];
      const double dbdt = planck_band_deriv(s2[DISORT_S_WN_LO], s2[DISORT_S_WN_HI], t);
      for (int f = 0; f < 6; ++f)
        for (int lu = 0; lu < ntau; ++lu)
          jac[i * per + ((int64_t)f * ntau + lu) * nd + k] = dbdt * flux[((size_t)field[f] * ne + e) * ntau + lu];
    }
  }
  return DISORT_OK;
}
